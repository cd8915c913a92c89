// CPU check of the wavefront plan of the temporally blocked 2D passes (tb2d_plan.h).
// Reads cases "nx ny xlo xhi ylo yhi src_ix drv_n" from stdin and, for depths 2..8 and
// several x chunks, checks that
//   - every owned cell of the grid belongs to exactly one wavefront of one class,
//   - every chunk lies in the grid and every rectangle's wave count matches its chunks,
//   - the slab-free bodies (MODE 0) load only planes and lanes inside 1..N-2 and outside
//     the slabs, the y-slab body (MODE 1) only such planes,
//   - a rectangle without excitation code has no source or drive plane within its marches.
//   - the y-slab and x-slab classes packed as one slab launch (set_slabs) hold exactly
//     their waves, each with its class's body (x-slab bit) and excitation flag.
// Prints one line per problem (at most a few per plan) and the totals.
#include <array>
#include <cstdio>
#include <vector>

#include "../lens_sensor_amd/csrc/tb2d_plan.h"

using namespace lfd;

static int check(const TBPlanIn &in, int T, int lx) {
    const TBPlan pl = tb_plan_2d(in, T, lx);
    const int Nx = in.nx, Ny = in.ny, W = 64 - 2 * T;
    std::vector<unsigned char> cov((size_t)Nx * Ny, 0);
    int bad = 0;
    auto err = [&](const char *m, int k, int g, int a, int b) {
        if (bad++ < 3)
            std::printf("BAD %s: %dx%d slabs x[%d,%d) y[%d,%d) src %d drv %d T %d lx %d class %d rect %d (%d, %d)\n",
                        m, Nx, Ny, in.xlo, in.xhi, in.ylo, in.yhi, in.src_ix, in.drv_n, T, lx, k, g, a, b);
    };
    for (int k = 0; k < kTBClasses; ++k) {
        const TBClass &q = pl.cls[k];
        if (q.nseg < 0 || q.nseg > 4) { err("rectangles", k, -1, q.nseg, 0); continue; }
        int prev = 0;
        for (int g = 0; g < q.nseg; ++g) {
            const int nw = q.end[g] - prev;
            prev = q.end[g];
            if (q.x0[g] < 0 || q.x1[g] > Nx || q.x1[g] <= q.x0[g] || q.lx[g] < 1) { err("range", k, g, q.x0[g], q.x1[g]); continue; }
            if (nw != (q.x1[g] - q.x0[g] + q.lx[g] - 1) / q.lx[g] * q.nwy[g]) err("count", k, g, nw, q.nwy[g]);
            const bool exc = q.exc == 1 || (q.exc == 2 && q.sexc[g]);
            for (int loc = 0; loc < nw; ++loc) {
                const int wy = q.wy0[g] + loc % q.nwy[g];
                const int X0 = q.x0[g] + (loc / q.nwy[g]) * q.lx[g];
                const int X1 = std::min(X0 + q.lx[g], q.x1[g]);
                const int j0 = wy * W - T;
                if ((q.mode == 0 || q.mode == 1) &&
                    (X0 - T < std::max(in.xlo, 1) || X1 + T - 1 > std::min(in.xhi, Nx - 2)))
                    err("x interior", k, g, X0, X1);
                if (q.mode == 0 && (j0 < std::max(in.ylo, 1) || j0 + 63 > std::min(in.yhi, Ny - 2)))
                    err("y interior", k, g, X0, wy);
                if ((q.mode & 2) && X1 + T - std::max(X0 - T, 0) > 64)     // x profiles in lanes
                    err("x-slab march longer than 64 planes", k, g, X0, X1);
                if (!exc) {
                    const int m0 = X0 - T, m1 = X1 + T;
                    if (in.src_ix >= 0 && in.src_ix < m1 && in.src_ix + 2 > m0) err("source", k, g, X0, X1);
                    if (in.drv_n > 0 && m1 > 0 && in.drv_n > m0) err("drive", k, g, X0, X1);
                }
                for (int x = X0; x < X1; ++x)
                    for (int l = T; l <= 63 - T; ++l) {
                        const int j = j0 + l;
                        if (j >= 0 && j < Ny) cov[(size_t)x * Ny + j]++;
                    }
            }
        }
    }
    for (size_t i = 0; i < cov.size(); ++i)
        if (cov[i] != 1) { err(cov[i] ? "covered twice" : "not covered", -1, -1, (int)(i / Ny), (int)(i % Ny)); break; }
    // the slab launch: every wave of classes 1 and 2, in order, with its class's body and
    // excitation flag, found the way the slab kernel finds it (segment, then chunk and wave)
    if (pl.cls[1].nseg + pl.cls[2].nseg > 0) {
        struct Segs {
            int nseg;
            int seg_x0[8], seg_x1[8], seg_wy0[8], seg_nwy[8], seg_end[8], seg_lx[8];
            unsigned seg_exc, seg_m3;
        } q{};
        set_slabs(q, pl.cls[1], pl.cls[2]);
        std::vector<std::array<int, 5>> want, got;
        for (int k : {1, 2}) {
            const TBClass &c = pl.cls[k];
            int prev = 0;
            for (int g = 0; g < c.nseg; ++g) {
                const bool exc = c.exc == 1 || (c.exc == 2 && c.sexc[g]);
                for (int loc = 0; loc < c.end[g] - prev; ++loc) {
                    const int X0 = c.x0[g] + (loc / c.nwy[g]) * c.lx[g];
                    want.push_back({k == 2, exc, X0, std::min(X0 + c.lx[g], c.x1[g]), c.wy0[g] + loc % c.nwy[g]});
                }
                prev = c.end[g];
            }
        }
        for (int wv = 0; wv < q.seg_end[q.nseg - 1]; ++wv) {
            int g = 0;
            while (g < q.nseg - 1 && wv >= q.seg_end[g]) ++g;
            const int loc = wv - (g > 0 ? q.seg_end[g - 1] : 0);
            const int X0 = q.seg_x0[g] + (loc / q.seg_nwy[g]) * q.seg_lx[g];
            got.push_back({(int)((q.seg_m3 >> g) & 1u), (int)((q.seg_exc >> g) & 1u), X0,
                           std::min(X0 + q.seg_lx[g], q.seg_x1[g]), q.seg_wy0[g] + loc % q.seg_nwy[g]});
        }
        if (q.nseg > 8 || got != want) err("slab launch", 12, q.nseg, (int)got.size(), (int)want.size());
    }
    return bad;
}

int main() {
    TBPlanIn in{};
    int total = 0, n = 0;
    while (std::scanf("%d %d %d %d %d %d %d %d", &in.nx, &in.ny, &in.xlo, &in.xhi, &in.ylo, &in.yhi,
                      &in.src_ix, &in.drv_n) == 8) {
        for (int T : {2, 3, 4, 6, 8})
            for (int lx : {3, 5, 16, 32, 64, 128, 1024}) {
                total += check(in, T, lx);
                ++n;
            }
    }
    std::printf("%d plans, %d problems\n", n, total);
    return total != 0;
}
