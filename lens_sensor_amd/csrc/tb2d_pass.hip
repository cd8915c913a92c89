// tb2d_pass.hip - the plain 2D blocked passes (no monitors).  Built once per polarisation
// with -DLF_TB2P_TM=<0|1>; each object defines its launcher, which lensfdtd.hip calls.
#include <hip/hip_runtime.h>

#include "tb2d.h"

#ifndef LF_TB2P_TM
#error "LF_TB2P_TM must be defined (0: TE, 1: TM)"
#endif

namespace lfd {

#if LF_TB2P_TM
using TB2Args = TBTMArgs;
template <int T, int MODE, bool EXC>
static void go(const TB2Args &q, const Excite &x, int fp64, hipStream_t st, unsigned nb) {
    tm_tb_kernel<T, MODE, EXC><<<nb, 256, 0, st>>>(q, x, fp64);
}
template <int T, int MODE>
static void go_mixed(const TB2Args &q, const Excite &x, int fp64, hipStream_t st, unsigned nb) {
    tm_tb_mixed_kernel<T, MODE><<<nb, 256, 0, st>>>(q, x, fp64);
}
#else
using TB2Args = TBTEArgs;
template <int T, int MODE, bool EXC>
static void go(const TB2Args &q, const Excite &x, int fp64, hipStream_t st, unsigned nb) {
    te_tb_kernel<T, MODE, EXC><<<nb, 256, 0, st>>>(q, x, fp64);
}
template <int T, int MODE>
static void go_mixed(const TB2Args &q, const Excite &x, int fp64, hipStream_t st, unsigned nb) {
    te_tb_mixed_kernel<T, MODE><<<nb, 256, 0, st>>>(q, x, fp64);
}
#endif

template <int T>
static void by_mode(int mode, int exc, const TB2Args &q, const Excite &x, int fp64, hipStream_t st,
                    unsigned nb) {
    if (mode == 0) (exc ? go<T, 0, true> : go<T, 0, false>)(q, x, fp64, st, nb);
    else if (mode == 1) (exc == 2 ? go_mixed<T, 1> : exc ? go<T, 1, true> : go<T, 1, false>)(q, x, fp64, st, nb);
    else (exc == 2 ? go_mixed<T, 3> : exc ? go<T, 3, true> : go<T, 3, false>)(q, x, fp64, st, nb);
}

#if LF_TB2P_TM
void launch_tm_tb(int T, int mode, int exc, const TBTMArgs &q, const Excite &x, int fp64,
                  hipStream_t st, unsigned nb) {
#else
void launch_te_tb(int T, int mode, int exc, const TBTEArgs &q, const Excite &x, int fp64,
                  hipStream_t st, unsigned nb) {
#endif
    if (T == 2) by_mode<2>(mode, exc, q, x, fp64, st, nb);
    else if (T == 4) by_mode<4>(mode, exc, q, x, fp64, st, nb);
    else if (T == 6) by_mode<6>(mode, exc, q, x, fp64, st, nb);
    else by_mode<8>(mode, exc, q, x, fp64, st, nb);
}

}  // namespace lfd
