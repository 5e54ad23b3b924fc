// Internal declarations shared by the kernels and the host runtime of
// libamrlbm.so. Not part of the C-ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "amrlbm.h"

namespace amrlbm {

// Block geometry of this build: 16 x 16 cells (the workloads' blocks).
constexpr int NX = 16;
constexpr int NY = 16;
constexpr int FW = NX + 2;            // population row width with the 1-cell ghost ring (LDS frame)
constexpr int FS = FW * (NY + 2);     // 324 cells per population per block
// One population array of a block in memory: the 256 interior cells as 4 x 4
// tiles of 4 x 4 cells (a tile is one 128-byte line, so a row or a column of
// the block touches 4 lines), then the 68 ghost-ring cells in ring order (see
// pcell), padded to a multiple of 128 bytes.
constexpr int GS = 336;
constexpr int POPB = 2 * 9 * GS;      // doubles per block per population buffer: h(9) then g(9)
constexpr int CW = NX + 4;            // C row width with 2 ghost rings
constexpr int CS = CW * (NY + 4);     // 400
constexpr int PS = NX * NY;           // 256
constexpr int PUB = 3 * PS;           // P, Ux, Uy

// Offset of cell (x, y), x, y in 0..17, inside a population array in memory.
// Ring order: left column y=1..16, right column, bottom row x=1..16, top row,
// then the corners (0,0), (0,17), (17,0), (17,17).
__host__ __device__ inline int pcell(int x, int y)
{
    if (x >= 1 && x <= NX && y >= 1 && y <= NY) {
        const int xi = x - 1, yi = y - 1;
        return (((yi >> 2) << 2) + (xi >> 2)) * 16 + ((yi & 3) << 2) + (xi & 3);
    }
    int t;
    if (y >= 1 && y <= NY)
        t = (x == 0 ? 0 : 16) + y - 1;
    else if (x >= 1 && x <= NX)
        t = (y == 0 ? 32 : 48) + x - 1;
    else
        t = 64 + (x == 0 ? 0 : 2) + (y == 0 ? 0 : 1);
    return NX * NY + t;
}

// Physics constants of D2Q9_Mod.f90 plus wc (D2Q9_Mod.f90:11).
struct Phys {
    double M0, wc, W, sigma, beta, kappa, taul, tauh, rhol, rhoh, gy;
    // quotients of the constants above, computed on the host with the same
    // IEEE division the kernels would use: (rhoh - rhol) / 3, (rhoh - rhol) / 6,
    // 1 / tauh, 1 / taul, and the correctly rounded reciprocal 1 / W that the
    // kernels' Markstein division by W takes
    double drho3, drho6, inv_tauh, inv_taul, inv_W;
};

// The same constants for the FP32 path, with the quotients it uses as
// products precomputed.
struct PhysF {
    float M0, wc, beta, kappa, taul, tauh, rhol, rhoh, gy;
    float drho, drho3, drho6, inv_tauh, inv_taul, k12W;   // k12W = 12 M0 / W
};

// Per-slot block descriptor on the device (slot = reference id - 1).
struct BlockDev {
    int level;    // -1 for free slots
    int dx;       // cell size in finest cells
    int I, J;     // block coordinates at its level
    int isleaf;
    int child0;   // slot of the first child when refined, else -1
    int hasleaf;  // children exist and are all leaves (derefinement candidate)
    int inst;     // instance of an ensemble mesh (0 in any other mesh)
};

// Neighbour descriptor: the resolved nghbr(k)%n plus what a ghost fill needs
// to know about that block, so one load serves the whole recipe.
struct NbDesc {
    int id;       // reference id (>0) or -k for a wall
    int level;
    int isleaf;
    int child0;   // slot of its first child when refined, else -1
};

// Device view of a mesh, passed by value to the kernels. T is the storage
// type of the fields: double (the reference's arithmetic) or float (the FP32
// path, which stores h - w_q and C - 1, deviations from the liquid at rest).
template <class T>
struct DevMeshT {
    T *pop[2];               // [slot][set 0:h 1:g][q][GS] (pcell order), ping-pong buffers
    T *C;                    // [slot][CS], C(-1:nx+2,-1:ny+2)
    T *PU;                   // [slot][3][PS]
    const BlockDev *blk;     // [slot]
    const NbDesc *nbd;       // [slot][8] neighbours in nghbr(1:8) order
    int *err_flags;          // [0]: bad level jump seen by a ghost fill
    const int *skip;         // [0] != 0: the step kernels return at once (a
                             // segment launched ahead of a regrid that changes
                             // the mesh)
    // an ensemble mesh (independent instances, each with its own fluid and
    // level range): per-instance constants indexed by BlockDev::inst, and
    // the roots per instance; null / 0 in any other mesh
    const Phys *phs;         // [inst]
    const PhysF *phsf;       // [inst]
    const int *lv;           // [inst][2]: min_levels, max_levels
    int tile_roots;
};

// The host's handle: the FP64 view, whose field pointers hold float arrays
// when f32 is set; the launchers pick the kernel instantiation.
struct DevMesh : DevMeshT<double> {
    int f32 = 0;
    PhysF phf{};
    int hg_depth = 1;   // hg fill operands in flight per thread: 0: 2/1, 1: 4/2, 2: 8/4 (first / later stages)
};

template <class T>
inline DevMeshT<T> typed(const DevMesh &m)
{
    DevMeshT<T> t;
    t.pop[0] = reinterpret_cast<T *>(m.pop[0]);
    t.pop[1] = reinterpret_cast<T *>(m.pop[1]);
    t.C = reinterpret_cast<T *>(m.C);
    t.PU = reinterpret_cast<T *>(m.PU);
    t.blk = m.blk;
    t.nbd = m.nbd;
    t.err_flags = m.err_flags;
    t.skip = m.skip;
    t.phs = m.phs;
    t.phsf = m.phsf;
    t.lv = m.lv;
    t.tile_roots = m.tile_roots;
    return t;
}

// Ghost program of one block for one fill kind: the block's own ghost cells
// and the ghost cells of coarser neighbours they read ("aux" cells), each one
// interpolation recipe over interior values or earlier cells, grouped into
// stages that depend only on earlier stages. Built on the host whenever the
// topology changes; evaluated in LDS by one workgroup per leaf.
enum FillKind { KIND_C1 = 0, KIND_C2 = 1, KIND_HG = 2 };
enum NodeType { NT_COPY = 0, NT_FINE = 1, NT_COARSE = 2, NT_BB_BOT = 3, NT_BB_TOP = 4 };
constexpr int MAX_STAGES = 6;
constexpr int NAUX_MAX = 128;       // aux cells per block and kind
// LDS cell strides: own cells first (C: 20x20, populations: 18x18), aux after
constexpr int LC_C = CS + NAUX_MAX;
constexpr int LC_F = FS + NAUX_MAX;
// most recipes in one block's program: its ghost cells plus NAUX_MAX aux cells
constexpr int NODES_HG = (FS - NX * NY) + NAUX_MAX;          // 1-ring of h, g
constexpr int NODES_C1 = (CS - NX * NY - 76) + NAUX_MAX;     // 1-ring of C (ring 2 not filled)
constexpr int NODES_C2 = (CS - NX * NY) + NAUX_MAX;          // 2 rings of C
struct __attribute__((aligned(16))) PNode {
    int dest;     // LDS cell written (own cell index < CS / FS, else aux)
    int info;     // type | wx pattern << 4 | wy pattern << 6
    int ref[9];   // >= 0: element offset in the global field; < 0: -(1 + LDS cell)
    int own;      // populations: pcell offset of an own ghost cell, -1 for aux cells
};
struct PHdr {
    int off;                  // first node
    int nst;                  // stages
    int end[MAX_STAGES];      // cumulative node count at the end of each stage
};
struct Prog {
    const PHdr *hdr;          // one per leaf, leaf-list order
    const PNode *nodes;
    int naux = 0;             // most aux cells of one leaf: a population's LDS stride is FS + naux
    int nmax = 0;             // most recipes of one leaf
};

// kernel launchers (amrlbm_kernels.hip); all enqueue on stream s
enum FillOp { OP_NONE = 0, OP_COLLIDE = 1, OP_MACRO_G = 2, OP_MACRO_G_COLLIDE = 3 };
// C fill (kind C1 or C2) from a ghost program, then op; with prog.hdr == null
// the block's ghosts are read from memory instead (no fill)
// learned regrid predictor (amrlbm_cnn.hip)
constexpr int CNN_NPARAM = 16 * 27 + 16 + 16 * 144 + 16 + 2 * 144 + 2;   // 3058
constexpr float kUScale = 10.0f;   // network input: velocity times this
void launch_cnn(const float *X, const int *idx, int n, const float *w, float *logits, const float *Y,
                float pos_weight, float *grad, float *loss, hipStream_t s);
void launch_cnn_adam(float *w, float *m, float *v, float *g, float lr, int t, hipStream_t s);
void launch_cnn_features(const DevMesh &m, const int *list, int n, float *X, double W, hipStream_t s);

// Control of one launch of the fused step's kernels: where the collision
// writes, and the device judge. The fused step collides out of place (the
// state before the step stays whole until the stream writes the next one), so
// a step the judge stops leaves the state of the step before it.
struct StepCtl {
    int out = -1;            // population buffer the collision writes (-1: in place, as the phase calls)
    int *judge = nullptr;    // the first ordinal (state after that many steps of the segment) at which the
                             // plain refinement criterion fired; none: INT_MAX. Null: no judge
    int ord = 0;             // this launch's state ordinal: the collision of step s and the stream of
                             // step s - 1 see the state after s steps
    int judge_on = 0;        // this launch tests the criterion on its state
    int min_level = 0, max_level = 0;   // the mesh's level range (an ensemble's per instance from m.lv)
    double refin_tol = 0.0;
};
void launch_prog_c(const DevMesh &m, int cur, const int *list, int n, int kind, int op, const Prog &prog,
                   const Phys &ph, hipStream_t s,
                   double *err_slot = nullptr, const StepCtl &ctl = StepCtl{});
// one int to v (a fill kernel, so that a captured graph holds no memset node)
void launch_fill_int(int *p, int v, hipStream_t s);
// out[kind][k] = slot[kind][list[k]] for k < n: the three kinds' ghost-program
// headers in launch order from those kept by block slot
void launch_gather_headers(PHdr *const slot[3], PHdr *const out[3], const int *list, int n, hipStream_t s);
// hg fill from a ghost program (or ghosts from memory with prog.hdr == null),
// then optionally the stream (+ C = sum h, + ghost ring carried over)
void launch_prog_hg(const DevMesh &m, int cur, const int *list, int n, const Prog &prog, bool stream_after,
                    bool macro_h, bool copy_ghosts, hipStream_t s, const int *judge = nullptr, int ord = 0);
void launch_macro_h(const DevMesh &m, int cur, const int *list, int n, hipStream_t s);
void launch_err(const DevMesh &m, const int *list, int n, double *err, double *err_slot, hipStream_t s);
// the device regrid decision and flags (k_regrid_check): at most RC_REF
// refinement flags and RC_CAND derefinement candidates in its lists
constexpr int RC_CAND = 512, RC_REF = 1024;
constexpr int kRegridDerefAt = 3 + RC_REF, kRegridFlagWords = 3 + RC_REF + RC_CAND;
constexpr size_t kRegridStaticLds = sizeof(int) * (18 * RC_CAND + RC_REF + 3);
int regrid_check_max_blocks();
hipError_t launch_regrid_check(const DevMesh &m, int nblks, const double *err_slot, int min_levels, int max_levels,
                               int levelmin, int levelmax, double refin_tol, double deref_tol, int *changed,
                               hipStream_t s, const int *judge = nullptr, int jstop = 0, int *flags = nullptr);
void launch_judge_check(const int *judge, int jstop, int *changed, hipStream_t s);
// pred[list[k]] = logits[2 k] > threshold and err_slot[list[k]] > gate (the
// learned predictor's refinement flags by block slot)
void launch_pred_flags(const float *logits, const int *list, int n, float threshold, const double *err_slot,
                       double gate, int *pred, hipStream_t s);
void launch_restrict(const DevMesh &m, int cur, const int *parents, int n, hipStream_t s);
void launch_prolong(const DevMesh &m, int cur, const int *pairs, int n, const Phys &ph, hipStream_t s);
void launch_init_fields(const DevMesh &m, int cur, const int *list, int n, const Phys &ph, double R, hipStream_t s,
                        const double *bub = nullptr, int nb = 0, const double *inst_R = nullptr);
// diagnostics of the leaves in list into out[9]; with ninst > 0 (an ensemble
// mesh) also per instance into out_inst[ninst][9]
void launch_diag(const DevMesh &m, const int *list, int n, const Phys &ph, double *partial, double *out,
                 hipStream_t s, int ninst = 0, double *out_inst = nullptr);
// Phase-field mass correction. launch_mass_sum: the FP64 sums of C dx^2 and
// of the interface weight max(0, 4 C (1 - C)) dx^2 over the leaves in list
// (part: 2 doubles per leaf), folded per instance (ninst workgroups) into
// mass[4 q .. 4 q + 3] = target, S, W, lambda as mode says; launch_mass_apply:
// dC = lambda max(0, 4 C (1 - C)) on every leaf cell, C += dC, h_q +=
// Gamma_q(U) dC, and g_q += (Gamma_q(U) - w_q) (rhoh - rhol) dC / 3, which
// keeps U and P.
// gate (sum only; apply always): the launch honours the skip word and the
// judge's stop ordinal ord, as the step kernels of a segment do.
enum MassMode { MASS_SAVE = 0, MASS_TARGET = 1, MASS_BUDGET = 2, MASS_CORRECT = 3 };
void launch_mass_sum(const DevMesh &m, const int *list, int n, int ninst, double *part, double *mass, int mode,
                     int gate, const int *judge, int ord, hipStream_t s);
void launch_mass_apply(const DevMesh &m, int cur, const int *list, int n, const double *mass, const Phys &ph,
                       const int *judge, int ord, hipStream_t s);
// n copies of 16-byte words: copy k takes desc[3k+2] words from src + desc[3k]
// to dst + desc[3k+1] (offsets in words)
// a segment of an upload batch: nwords 4-byte words from word src_word of the
// batch to the device address dst
struct ScatterSeg {
    unsigned long long dst;
    int src_word, nwords;
};
void launch_scatter_segments(const void *src, const ScatterSeg *desc, int n, hipStream_t s);

// The persistent segment kernel: len whole steps of every leaf in list (one
// workgroup each; the caller checks with segment_capacity that all fit on the
// GPU at once) in one launch, with the same results as the two-kernel step
// graph of the segment. sync: the grid barrier's words (kSegSyncWords ints,
// zeroed by the launch except word 15, the fault word, set when a barrier
// wait timed out). Returns the launch's status.
constexpr int kSegSyncWords = 16 + 16 * 8;
hipError_t launch_segment(const DevMesh &m, int cur, const int *list, int n, const Prog &c2, const Prog &hg,
                          const Prog &c1, const Phys &ph, int len, double *err_slot, unsigned *sync, hipStream_t s);
// dynamic LDS of one segment workgroup, and how many the GPU holds at once
size_t segment_lds_bytes(const DevMesh &m, const Prog &hg);
int segment_capacity(const DevMesh &m, const Prog &hg);

}  // namespace amrlbm
