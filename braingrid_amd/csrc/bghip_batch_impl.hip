// Multi-network launcher of the implicit-ring fused kernel (k_fused_impl_batch in
// bghip_implicit.h), in a translation unit of its own so that it compiles in parallel.
#include "bghip_implicit.h"
#include "bghip_launch.h"
#include <algorithm>

template <int NK, int SK>
static void fused_impl(const BatchMap &bm, hipStream_t s, int blk, int kmax, int K, int Kfull, int launch)
{
    (void)blk;   // one workgroup of BG_BLOCK threads per four destinations of a member
    const dim3 grid(std::max<uint32_t>(bm.nquads, 1)), block(BG_BLOCK);
    if (K == BG_K_COMMON)
        hipLaunchKernelGGL((k_fused_impl_batch<NK, SK, BG_K_COMMON>), grid, block, 0, s, bm, K, Kfull, launch);
    else if (K < BG_K_COMMON)
        hipLaunchKernelGGL((k_fused_impl_batch<NK, SK, 8>), grid, block, 0, s, bm, K, Kfull, launch);
    else if (kmax <= 16)
        hipLaunchKernelGGL((k_fused_impl_batch<NK, SK, 16>), grid, block, 0, s, bm, K, Kfull, launch);
    else
        hipLaunchKernelGGL((k_fused_impl_batch<NK, SK, 32>), grid, block, 0, s, bm, K, Kfull, launch);
}

void launch_fused_impl_batch(const BatchMap &bm, hipStream_t s, const BatchShape &sh, int nk, int sk, int kmax,
                             int K, int Kfull, int launch)
{
    if (nk == NK_LIF) {
        if (sk == SK_DS) fused_impl<NK_LIF, SK_DS>(bm, s, sh.blk, kmax, K, Kfull, launch);
        else fused_impl<NK_LIF, SK_SPIKING>(bm, s, sh.blk, kmax, K, Kfull, launch);
    } else {
        if (sk == SK_DS) fused_impl<NK_IZH, SK_DS>(bm, s, sh.blk, kmax, K, Kfull, launch);
        else fused_impl<NK_IZH, SK_SPIKING>(bm, s, sh.blk, kmax, K, Kfull, launch);
    }
}

