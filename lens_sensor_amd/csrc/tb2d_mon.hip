// tb2d_mon.hip - monitored 2D blocked passes (the plain pass plus the CW-window monitors
// summed inside it).  Built once per polarisation with -DLF_TB2M_TM=<0|1>; each object
// defines its launcher, which lensfdtd.hip calls.
#include <hip/hip_runtime.h>

#include "tb2d.h"

#ifndef LF_TB2M_TM
#error "LF_TB2M_TM must be defined (0: TE, 1: TM)"
#endif

namespace lfd {

#if LF_TB2M_TM
using TB2Args = TBTMArgs;
template <int T, int MODE, bool EXC, int MON>
static void go(const TB2Args &q, const Excite &x, const TBMon &m, int fp64, hipStream_t st, unsigned nb) {
    tm_tb_mon_kernel<T, MODE, EXC, MON><<<nb, 256, 0, st>>>(q, x, m, fp64);
}
#else
using TB2Args = TBTEArgs;
template <int T, int MODE, bool EXC, int MON>
static void go(const TB2Args &q, const Excite &x, const TBMon &m, int fp64, hipStream_t st, unsigned nb) {
    te_tb_mon_kernel<T, MODE, EXC, MON><<<nb, 256, 0, st>>>(q, x, m, fp64);
}
#endif

template <int T, int MON>
static void by_mode(int mode, bool exc, const TB2Args &q, const Excite &x, const TBMon &m, int fp64,
                    hipStream_t st, unsigned nb) {
    if (mode == 0 && !exc) go<T, 0, false, MON>(q, x, m, fp64, st, nb);
    else if (mode == 0) go<T, 0, true, MON>(q, x, m, fp64, st, nb);
    else if (mode == 1) go<T, 1, true, MON>(q, x, m, fp64, st, nb);
    else go<T, 3, true, MON>(q, x, m, fp64, st, nb);
}

template <int T>
static void by_mon(int mon, int mode, bool exc, const TB2Args &q, const Excite &x, const TBMon &m,
                   int fp64, hipStream_t st, unsigned nb) {
    // mon bit 4: every phasor accumulates in complex128 (no per-level format test);
    // bit 5 (TM only): the fixed CW-window layout of the stitch runs
    if (mon == 9) by_mode<T, 9>(mode, exc, q, x, m, fp64, st, nb);
    else if (mon == 14) by_mode<T, 14>(mode, exc, q, x, m, fp64, st, nb);
    else if (mon == 25) by_mode<T, 25>(mode, exc, q, x, m, fp64, st, nb);
#if LF_TB2M_TM
    else if (mon == 62) by_mode<T, 62>(mode, exc, q, x, m, fp64, st, nb);
#endif
    else by_mode<T, 30>(mode, exc, q, x, m, fp64, st, nb);
}

#if LF_TB2M_TM
void launch_tm_tb_mon(int T, int mode, bool exc, int mon, const TBTMArgs &q, const Excite &x,
                      const TBMon &m, int fp64, hipStream_t st, unsigned nb) {
#else
void launch_te_tb_mon(int T, int mode, bool exc, int mon, const TBTEArgs &q, const Excite &x,
                      const TBMon &m, int fp64, hipStream_t st, unsigned nb) {
#endif
    if (T == 2) by_mon<2>(mon, mode, exc, q, x, m, fp64, st, nb);
    else if (T == 3) by_mon<3>(mon, mode, exc, q, x, m, fp64, st, nb);
    else by_mon<4>(mon, mode, exc, q, x, m, fp64, st, nb);
}

}  // namespace lfd
