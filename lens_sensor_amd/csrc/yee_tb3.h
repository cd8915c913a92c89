// yee_tb3.h - the temporally blocked 3D pass: tile boxes, launch arguments and the
// launchers compiled in yee_tb3.hip (one object per depth and loss variant).
#pragma once
#include "lensfdtd_dev.h"

namespace lfd {

// ---------------------------------------------------------------------------- temporal blocking (3D)
// One workgroup of kTB3W wavefronts advances one tile of the grid by T steps per pass.
// Lane l of every wavefront holds the z cell k = K0 - T + l, wavefront w holds the R rows
// j = J0 - T + w*R + r, and the workgroup marches along x from X0 - T to X1 + T - 1.
// While plane p is loaded, level s (step n+s -> n+s+1) advances plane q = p-1-s: H(q)
// from the level's inputs E(q) (loaded / produced one iteration earlier) and E(q+1)
// (this iteration's output of level s-1), then E(q) from H(q) and the level's own H(q-1).
// z neighbours come from DPP lane shifts, y neighbours from the adjacent row in registers
// or, across wavefronts, from LDS (one workgroup barrier per level).  Every level loses
// one cell of validity per side along y and z (and the first plane along x), so after T
// levels exactly the owned cells K0..K1-1, J0..J1-1, X0..X1-1 hold step n+T.  Fields and
// every psi slab are read from the step-n buffers and written to the other buffers, so a
// recomputed halo value never sees a partly advanced neighbour.  Per cell and step the
// arithmetic is that of yee_h_kernel + yee_e_kernel (boundaries3d.py:311-362).
constexpr int kTB3W = 16;            // wavefronts per workgroup
constexpr int kTB3MaxBox = 128;
constexpr int kTB3BoxSlots = 16384;   // per-tile FP64 slots of a fused detector box

struct TB3Box {                      // a box of tiles of one wavefront class
    int mode, rows;                  // class bits (1 z, 2 y, 4 x: slab or edge), rows per wavefront
    int exc;                         // the tiles' march meets a source or drive plane
    int box;                         // the tiles' planes meet a fused detector box
    int x0, x1, lx;                  // owned x range, x chunk length
    int y0, y1, z0, z1;              // owned y / z ranges
    int nty, ntz;                    // tiles along y and z (x chunks make up the rest)
    int first;                       // index of the box's first tile in the pass
};

struct TB3Args {
    Grid g;
    const float *Ex0, *Ey0, *Ez0, *Hx0, *Hy0, *Hz0;
    float *Ex1, *Ey1, *Ez1, *Hx1, *Hy1, *Hz1;
    const float *cex, *cey, *cez, *dax, *day, *daz;
    float ch;
    Prof pe[3], ph[3];
    // psi slabs, step-n (s0) and step-n+T (s1) buffers: 0 Hx_z 1 Hx_y 2 Hy_x 3 Hy_z
    // 4 Hz_y 5 Hz_x 6 Ex_y 7 Ex_z 8 Ey_z 9 Ey_x 10 Ez_x 11 Ez_y
    Strip s0[12], s1[12];
    int pec[6];
    const TB3Box *box;               // box table on the device
    int nbox, ntile;                 // boxes, tiles of the pass
    unsigned *ticket;                // tile tickets: workgroups draw tiles until ntile
    unsigned base;                   // ticket value of this pass's first tile
    StepArgs st[4];
    unsigned mask[4];                // monitor mask of each step of the pass
    // detector boxes summed inside the pass (sum of Ex^2 + Ey^2 + Ez^2 over a box, every
    // step whose mask has the box's bit): per box and component (Ex, Ey, Ez) the index
    // ranges, and one FP64 slot per tile
    int nboxm;
    unsigned boxbit[2];
    int blo[2][3][3], bhi[2][3][3];
    double *bpart[2];
};

// launch of one blocked pass: depth T (1, 2 or 3), loss factors present or not
void launch_yee_tb3(int T, bool loss, const TB3Args &f, const Excite &x, int fp64, unsigned nb,
                    hipStream_t st);
void launch_yee_tb3_1_0(const TB3Args &, const Excite &, int, unsigned, hipStream_t);
void launch_yee_tb3_1_1(const TB3Args &, const Excite &, int, unsigned, hipStream_t);
void launch_yee_tb3_2_0(const TB3Args &, const Excite &, int, unsigned, hipStream_t);
void launch_yee_tb3_2_1(const TB3Args &, const Excite &, int, unsigned, hipStream_t);
void launch_yee_tb3_3_0(const TB3Args &, const Excite &, int, unsigned, hipStream_t);
void launch_yee_tb3_3_1(const TB3Args &, const Excite &, int, unsigned, hipStream_t);

}  // namespace lfd
