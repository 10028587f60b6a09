// The one-pass VALU level with its range decimation fixed at 2 (the later levels of many plans: their loops over
// phases exact, no reads or selects for the phases a run-time factor would skip), instantiated in a translation unit
// of its own so that it compiles beside the run-time form. launch_level_fused takes it where Dk = 2.
#include "fsh_common.h"
#include "fsh_kernels.h"
#include "fused_level.h"

namespace fsh {

int launch_level_fused_d2(dim3 grid, size_t lds, hipStream_t s, int NWK, int NWP, const float2* x, int P, int K,
                          const float* c0, const float* sl, int C, const float* hk_t, int plk, int Ko, const float* hp_t,
                          int Dp, int plp, int Po, double kc, int Np, int CB, int R, float2* z)
{
#define FSH_LF(A, B) hipLaunchKernelGGL((k_level_fused<A, B, 2>), grid, dim3(256), lds, s, x, P, K, c0, sl, C, hk_t, 2, plk, \
                                        Ko, hp_t, Dp, plp, Po, kc, Np, CB, R, z)
#define FSH_LF_P(A)                                                                                                   \
    switch (NWP) {                                                                                                    \
        case 1: FSH_LF(A, 1); break;  case 2: FSH_LF(A, 2); break;  case 3: FSH_LF(A, 3); break;                     \
        case 4: FSH_LF(A, 4); break;  case 6: FSH_LF(A, 6); break;  case 8: FSH_LF(A, 8); break;                     \
        case 9: FSH_LF(A, 9); break;  case 10: FSH_LF(A, 10); break; case 12: FSH_LF(A, 12); break;                  \
        default: FSH_LF(A, 16); break;                                                                                \
    }
    switch (NWK) {
        case 1: FSH_LF_P(1); break;   case 2: FSH_LF_P(2); break;   case 3: FSH_LF_P(3); break;
        case 4: FSH_LF_P(4); break;   case 6: FSH_LF_P(6); break;   case 8: FSH_LF_P(8); break;
        case 9: FSH_LF_P(9); break;   case 10: FSH_LF_P(10); break; case 12: FSH_LF_P(12); break;
        default: FSH_LF_P(16); break;
    }
#undef FSH_LF_P
#undef FSH_LF
    FSH_TRY(hipGetLastError());
    return OK;
}

}  // namespace fsh
