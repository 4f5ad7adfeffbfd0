// Implicit-ring fused block kernels for Izhikevich neurons (k_fused_impl in bghip_implicit.h),
// in a translation unit of their own so that they compile in parallel with the explicit ones.
#include "bghip_implicit.h"
#include "bghip_launch.h"

// implicit-ring blocks (k_fused_impl in bghip_implicit.h): 32-bit rings only
template <int SK>
static void go_impl(const DevNet &dn, hipStream_t s, int kmax, uint64_t t0, int K, int sel)
{
    unsigned blocks = (unsigned)(((uint64_t)dn.N * BG_WAVE + BG_BLOCK - 1) / BG_BLOCK);
    if (K == BG_K_COMMON)
        hipLaunchKernelGGL((k_fused_impl<NK_IZH, SK, BG_K_COMMON>), dim3(blocks), dim3(BG_BLOCK), 0, s, dn, t0, K, sel);
    else if (K < BG_K_COMMON)
        hipLaunchKernelGGL((k_fused_impl<NK_IZH, SK, 8>), dim3(blocks), dim3(BG_BLOCK), 0, s, dn, t0, K, sel);
    else if (kmax <= 16)
        hipLaunchKernelGGL((k_fused_impl<NK_IZH, SK, 16>), dim3(blocks), dim3(BG_BLOCK), 0, s, dn, t0, K, sel);
    else
        hipLaunchKernelGGL((k_fused_impl<NK_IZH, SK, 32>), dim3(blocks), dim3(BG_BLOCK), 0, s, dn, t0, K, sel);
}

void launch_fused_impl_izh(const DevNet &dn, hipStream_t s, int sk, int kmax, uint64_t t0, int K, int sel)
{
    if (sk == SK_DS) go_impl<SK_DS>(dn, s, kmax, t0, K, sel);
    else go_impl<SK_SPIKING>(dn, s, kmax, t0, K, sel);
}
