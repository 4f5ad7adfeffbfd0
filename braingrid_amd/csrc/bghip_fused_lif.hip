// Fused block kernels for LIF neurons (see k_fused_block in bghip_kernels.h).
#include "bghip_kernels.h"
#include "bghip_launch.h"

template <int SK, int W32>
static void go(const DevNet &dn, hipStream_t s, int kmax, uint64_t t0, int K, int pull, uint64_t tprev0, int sel)
{
    unsigned blocks = (unsigned)(((uint64_t)dn.N * BG_WAVE + BG_BLOCK - 1) / BG_BLOCK);
    if (W32 == 1 && K == BG_K_COMMON)   // the usual full block of a 32-bit-ring network
        hipLaunchKernelGGL((k_fused_block<NK_LIF, SK, W32, BG_K_COMMON>), dim3(blocks), dim3(BG_BLOCK), 0, s, dn, t0, K,
                           pull, tprev0, sel);
    else if (kmax <= 16)
        hipLaunchKernelGGL((k_fused_block<NK_LIF, SK, W32, 16>), dim3(blocks), dim3(BG_BLOCK), 0, s, dn, t0, K, pull,
                           tprev0, sel);
    else
        hipLaunchKernelGGL((k_fused_block<NK_LIF, SK, W32, 32>), dim3(blocks), dim3(BG_BLOCK), 0, s, dn, t0, K, pull,
                           tprev0, sel);
}

template <int SK>
static void go_w(const DevNet &dn, hipStream_t s, int kmax, uint64_t t0, int K, int pull, uint64_t tprev0, int sel)
{
    switch (dn.W32) {
    case 1: go<SK, 1>(dn, s, kmax, t0, K, pull, tprev0, sel); break;
    case 2: go<SK, 2>(dn, s, kmax, t0, K, pull, tprev0, sel); break;
    case 4: go<SK, 4>(dn, s, kmax, t0, K, pull, tprev0, sel); break;
    default: go<SK, 8>(dn, s, kmax, t0, K, pull, tprev0, sel); break;
    }
}

void launch_fused_lif(const DevNet &dn, hipStream_t s, int sk, int kmax, uint64_t t0, int K, int pull,
                      uint64_t tprev0, int sel)
{
    if (sk == SK_DS) go_w<SK_DS>(dn, s, kmax, t0, K, pull, tprev0, sel);
    else go_w<SK_SPIKING>(dn, s, kmax, t0, K, pull, tprev0, sel);
}
