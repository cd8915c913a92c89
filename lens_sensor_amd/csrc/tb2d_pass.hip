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
template <int T, int MODE, int EXC>
static void go(const TB2Args &q, const Excite &x, int fp64, hipStream_t st, unsigned nb) {
    tm_tb_kernel<T, MODE, EXC><<<nb, 256, 0, st>>>(q, x, fp64);
}
template <int T, int MODE, int EXC>
static void go_mixed(const TB2Args &q, const Excite &x, int fp64, hipStream_t st, unsigned nb) {
    tm_tb_mixed_kernel<T, MODE, EXC><<<nb, 256, 0, st>>>(q, x, fp64);
}
template <int T, int EXC>
static void go_slab(const TB2Args &q, const Excite &x, int fp64, hipStream_t st, unsigned nb) {
    tm_tb_slab_kernel<T, EXC><<<nb, 256, 0, st>>>(q, x, fp64);
}
#else
using TB2Args = TBTEArgs;
template <int T, int MODE, int EXC>
static void go(const TB2Args &q, const Excite &x, int fp64, hipStream_t st, unsigned nb) {
    te_tb_kernel<T, MODE, EXC><<<nb, 256, 0, st>>>(q, x, fp64);
}
template <int T, int MODE, int EXC>
static void go_mixed(const TB2Args &q, const Excite &x, int fp64, hipStream_t st, unsigned nb) {
    te_tb_mixed_kernel<T, MODE, EXC><<<nb, 256, 0, st>>>(q, x, fp64);
}
template <int T, int EXC>
static void go_slab(const TB2Args &q, const Excite &x, int fp64, hipStream_t st, unsigned nb) {
    te_tb_slab_kernel<T, EXC><<<nb, 256, 0, st>>>(q, x, fp64);
}
#endif

// exc: the class's rectangles with excitation (0 none, 1 all, 2 some); em: the code they
// need: 1 source only, 4 drive only with the float64 blend (the 2D stitch windows after
// the first), 3 anything else (source and drive, blend chosen at run time)
template <int T, int MODE, int EM>
static void by_exc(int exc, const TB2Args &q, const Excite &x, int fp64, hipStream_t st, unsigned nb) {
    if (exc == 0) go<T, MODE, 0>(q, x, fp64, st, nb);
    else if constexpr (MODE == 0) go<T, MODE, EM>(q, x, fp64, st, nb);   // no mixed slab-free class
    else if (exc == 1) go<T, MODE, EM>(q, x, fp64, st, nb);
    else go_mixed<T, MODE, EM>(q, x, fp64, st, nb);
}

template <int T, int MODE>
static void by_em(int exc, int em, const TB2Args &q, const Excite &x, int fp64, hipStream_t st,
                  unsigned nb) {
    if (em == 1) by_exc<T, MODE, 1>(exc, q, x, fp64, st, nb);
    else if (em == 4) by_exc<T, MODE, 4>(exc, q, x, fp64, st, nb);
    else by_exc<T, MODE, 3>(exc, q, x, fp64, st, nb);
}

template <int T>
static void by_mode(int mode, int exc, int em, const TB2Args &q, const Excite &x, int fp64,
                    hipStream_t st, unsigned nb) {
    if (mode == 0) by_em<T, 0>(exc, em, q, x, fp64, st, nb);
    else if (mode == 1) by_em<T, 1>(exc, em, q, x, fp64, st, nb);
    else by_em<T, 3>(exc, em, q, x, fp64, st, nb);
}

// excitation code a context's passes need: 0 none, 1 source only, 4 drive only with the
// float64 blend, 3 anything else (a source or drive that the stepped component does not
// carry is never applied)
static int excitation_code(const Excite &x) {
#if LF_TB2P_TM
    const int comp = LF_EY;
#else
    const int comp = LF_EZ;
#endif
    const bool src = x.amp != nullptr && x.src_comp == comp;
    const bool drv = x.drv[comp] != nullptr && x.drv_n > 0;
    return !src && !drv ? 0 : !drv ? 1 : (!src && x.drv_math == LF_DRIVE_C128) ? 4 : 3;
}

// the host merges the slab classes only without excitation or with the drive alone (their
// bodies then hold two waves per SIMD together); the generic body serves any other case
template <int T>
static void slab_by_em(int em, const TB2Args &q, const Excite &x, int fp64, hipStream_t st, unsigned nb) {
    if (em == 0) go_slab<T, 0>(q, x, fp64, st, nb);
    else if (em == 4) go_slab<T, 4>(q, x, fp64, st, nb);
    else go_slab<T, 3>(q, x, fp64, st, nb);
}

#if LF_TB2P_TM
void launch_tm_tb_slab(int T, int exc, const TBTMArgs &q, const Excite &x, int fp64, hipStream_t st,
                       unsigned nb) {
#else
void launch_te_tb_slab(int T, int exc, const TBTEArgs &q, const Excite &x, int fp64, hipStream_t st,
                       unsigned nb) {
#endif
    const int em = exc ? excitation_code(x) : 0;
    if (T == 2) slab_by_em<2>(em, q, x, fp64, st, nb);
    else if (T == 4) slab_by_em<4>(em, q, x, fp64, st, nb);
    else if (T == 6) slab_by_em<6>(em, q, x, fp64, st, nb);
    else slab_by_em<8>(em, q, x, fp64, st, nb);
}

#if LF_TB2P_TM
void launch_tm_tb(int T, int mode, int exc, const TBTMArgs &q, const Excite &x, int fp64,
                  hipStream_t st, unsigned nb) {
#else
void launch_te_tb(int T, int mode, int exc, const TBTEArgs &q, const Excite &x, int fp64,
                  hipStream_t st, unsigned nb) {
#endif
    const int em = excitation_code(x);
    if (em == 0) exc = 0;
    if (T == 2) by_mode<2>(mode, exc, em, q, x, fp64, st, nb);
    else if (T == 4) by_mode<4>(mode, exc, em, q, x, fp64, st, nb);
    else if (T == 6) by_mode<6>(mode, exc, em, q, x, fp64, st, nb);
    else by_mode<8>(mode, exc, em, q, x, fp64, st, nb);
}

}  // namespace lfd
