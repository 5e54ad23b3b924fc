// Host runtime of libamrlbm.so: mesh topology, regrid bookkeeping, device
// tables and the C-ABI declared in include/amrlbm.h.
//
// Topology model. The reference stores each neighbour as a Fortran pointer
// into Blk(q)%child(c) of the parent's same-level neighbour q
// (Neighbors.f90:19-37), so it reads as q's child c once q is refined and as q
// itself while q is a leaf. Under 2:1 balance that is exactly "the block at
// the adjacent position on my level if it exists, else the block covering it
// one level up", which this runtime evaluates from a per-level position map.
// Wall directions read -k (AMR_Rising_2D_OMP.f90:105, Neighbors.f90:47-50).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <omp.h>
#include <parallel/algorithm>
#include <vector>

#include "amrlbm_internal.h"

using namespace amrlbm;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

#define HIPCHK(expr)                                                                                \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return fail(AMRLBM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));         \
    } while (0)

#define NEED_READY(m)                                                       \
    do {                                                                    \
        if (!(m)) return fail(AMRLBM_ERR_ARG, "null mesh");                 \
        if (!(m)->ready) return fail(AMRLBM_ERR_STATE, "mesh not set up"); \
    } while (0)

struct Block {
    int level = -1, dx = 0, I = 0, J = 0, parent = 0;
    int child[4] = {0, 0, 0, 0};
    bool has_child = false, isleaf = false, hasleaf = false, refin = false, deref = false;
    // the reference's Blk(n)%h etc. are ALLOCATED: from Add_Child_Blocks
    // (AMR_Rising_2D_OMP.f90:565-570) until Prolong frees a refined block's
    // (Prolong.f90:26), or Restrict allocates a derefined one's
    // (Restrict.f90:25-27) until Remove_Children frees its children's
    // (Refine_Derefine.f90:242-243). Initial_Mesh's inner blocks keep theirs,
    // zero. Restart files carry arrays for exactly these blocks.
    bool alloc = false;
};

// Ghost program of one leaf for one fill kind, nodes grouped by stage.
struct LeafProg {
    std::vector<PNode> nodes;
    int naux = 0;
    int nst = 0;
    int end[MAX_STAGES] = {};
    std::string err;
};

// what a program build reads of a block
struct BlockSig {
    int level, I, J, child0, flags;
    bool operator==(const BlockSig &o) const
    {
        return level == o.level && I == o.I && J == o.J && child0 == o.child0 && flags == o.flags;
    }
};
inline BlockSig block_sig(const Block &b)
{
    return {b.level, b.I, b.J, b.child[0], (b.isleaf ? 1 : 0) | (b.has_child ? 2 : 0)};
}

// The ghost programs of one block id and what they were built from: every
// block the builds read (neighbours, their children, coarse neighbours'
// neighbours), with its level, position, first child and leaf flag then. The
// builds read nothing else of the topology, so while all of these are equal
// the programs are too and are reused after a regrid.
struct ProgCache {
    LeafProg p[3];
    std::vector<std::pair<int, BlockSig>> deps;
    bool valid = false;
    bool on_device = false;   // p is in the block's slot of the device arenas
};

}  // namespace

struct amrlbm_mesh {
    amrlbm_params prm{};
    Phys ph{};
    std::vector<Block> blk;          // 1-based, size max_blocks + 1
    int nblks = 0, kk = 0, T = 0, levelmin = 0, levelmax = 0;
    int tile_roots = 0;              // roots per independent instance (walls between instances)
    std::deque<int> freeq;
    std::vector<int> leaf;           // Array_of_Leaf_Blocks order (ascending id)
    std::vector<int> disp;           // the leaves in launch order (see dispatch_order)
    std::vector<std::vector<int>> at;  // per level: position -> id (0 = none)
    bool ready = false;

    // device state
    hipStream_t stream = nullptr;
    DevMesh dm{};
    BlockDev *d_blk = nullptr;
    NbDesc *d_nbd = nullptr;
    int *d_err_flags = nullptr;      // [0] a fill met a level jump (d_changed[3]: it travels with each decision)
    unsigned *d_sync = nullptr;      // the segment kernel's grid barrier words (fault word at 15)
    int *d_judge = nullptr;          // the device judge's stop ordinal (enqueue_segment with judge)
    // the learned regrid predictor of amrlbm_step_learned: its weights, the
    // network inputs and logits of every leaf, its refinement flags by slot
    float *d_net_w = nullptr, *d_net_x = nullptr, *d_net_logit = nullptr;
    int *d_pred = nullptr;
    int net_cap = 0;
    bool has_net = false;
    float net_threshold = 0.0f;
    double net_gate = 0.25;          // Err_Interface a predicted block must reach, in refinement tolerances
    int *d_lists = nullptr;          // the leaves in launch order (slot indices, id - 1)
    int lists_cap = 0;
    int *d_pairs = nullptr;          // restrict / prolong work lists
    int pairs_cap = 0;
    double *d_err = nullptr;         // Err_Interface per leaf
    double *d_err_slot = nullptr;    // Err_Interface per block slot (leaves)
    int *d_changed = nullptr;        // [0] the device regrid decision (the step kernels' skip word),
                                     // [1] any refinement flag, [2] a judged segment's stop ordinal,
                                     // [3] the error word d_err_flags,
                                     // [4 ..] the decision's flags (k_regrid_check's flag lists)
    int *h_changed = nullptr;        // pinned copy of it
    double *d_part = nullptr;        // diagnostics partials
    double *d_diag = nullptr;
    double *h_pinned = nullptr;      // staging (err, diagnostics)
    double *scratch = nullptr;       // this device's copy of the file-scope scratch array
    int cur = 0;
    bool tables_dirty = true;
    // the host's copy of the device block and neighbour tables; between two
    // uploads, the blocks whose rows a regrid changed (rows_all: every row,
    // after a set-up, a restart or a caller-built topology)
    std::vector<BlockDev> h_bd;
    std::vector<NbDesc> h_nbd;
    std::vector<int> dirty_rows;
    std::vector<char> dirty_mark;
    bool rows_all = true;
    // the ids whose rows the last upload rebuilt, ascending (every id whose
    // leaf status changed is among them); leaf_delta_all: every id
    std::vector<int> leaf_delta;
    bool leaf_delta_all = true;
    bool leaf_list_bad = false;      // AMRLBM_VERIFY_TABLES: a merged leaf list differed from a scan
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_check = nullptr;
    hipEvent_t ev_tab = nullptr;   // after the last batch scatter (and the work before it)
    bool tab_recorded = false;     // ev_tab has been recorded (an unrecorded event must not be waited on)
    hipEvent_t ev_pairs = nullptr; // after the last copy out of h_pairs
    bool pairs_recorded = false;
    double last_step_ms = 0.0;
    int regrid_checks = 0, regrid_changes = 0, segments_skipped = 0, graphs_built = 0;
    bool seg_kernel_ran = false;     // the persistent segment kernel ran in the current stepping call
    int device_flag_regrids = 0;     // regrids whose flags came from the device decision
    // host wall time (ms): Refine_Derefine_Mesh, table upload, ghost-program
    // build (part of the upload), graph capture and instantiation
    double host_ms_regrid = 0.0, host_ms_tables = 0.0, host_ms_programs = 0.0, host_ms_graphs = 0.0;

    // ghost programs per fill kind (C1, C2, HG), leaf-list order
    // the programs on the device: per kind, block id n's nodes at
    // (n - 1) * kProgSlot[kind] of d_pnode (an arena of prog_slots slots),
    // so a mesh change uploads only the leaves whose programs were rebuilt;
    // per leaf (launch order) a header pointing there
    PHdr *d_phdr[3] = {nullptr, nullptr, nullptr};        // per leaf, launch order (gathered on the device)
    PHdr *d_phdr_slot[3] = {nullptr, nullptr, nullptr};   // per block slot, as built
    PNode *d_pnode[3] = {nullptr, nullptr, nullptr};
    size_t phdr_cap[3] = {0, 0, 0};
    int prog_slots = 0;
    // The upload batch of a table upload: the block and neighbour tables, the
    // leaf lists, the rebuilt ghost programs and the program headers, packed
    // into one pinned buffer, sent with one copy and put in place by one
    // scatter launch (segments listed in segs)
    char *h_up = nullptr;
    size_t up_cap = 0, up_used = 0;
    void *d_up = nullptr;
    size_t dup_cap = 0;
    std::vector<ScatterSeg> segs;
    int *h_pairs = nullptr;   // pinned restrict / prolong work lists
    size_t hpairs_cap = 0;
    long long prog_nodes_uploaded = 0;
    Prog prog[3] = {};
    std::vector<ProgCache> pcache;   // by block id
    // by block id: the ids whose cached programs read the block (ProgCache
    // deps); the blocks whose table rows the last upload changed, whose
    // dependants are the only programs that can be stale (check_all: every
    // leaf's, after a full upload)
    std::vector<std::vector<int>> rdeps;
    std::vector<int> changed_blocks;
    bool check_all = true;
    // the ids whose Refin or Deref flag the last flag pass set
    std::vector<int> flagged;
    // the leaves' Morton keys, sorted, kept across mesh changes (orders 0, 1,
    // 3): the blocks whose rows changed since are re-keyed and merged in
    std::vector<std::pair<unsigned long long, int>> morton, morton_next;
    std::vector<int> morton_changed;
    std::vector<char> morton_drop;
    bool morton_all = true;
    // per kind and block id: the header of its cached program (offset into its
    // slot), its aux cells and recipes, so a mesh change gathers the headers of
    // every leaf from compact arrays
    std::vector<PHdr> hdr_id[3];
    std::vector<int> naux_id[3], nn_id[3];
    std::vector<int> cost_id;   // operand loads of the C2 and hg programs (launch order 3)
    // ids whose programs were rebuilt since the last program upload
    std::vector<int> prog_pending;
    // per kind, over the leaves: how many programs have each aux-cell count
    // and each recipe count (the kernels size their LDS by the largest);
    // counted: the id's naux_id / nn_id are in the counts
    std::vector<int> hist_aux[3], hist_nn[3];
    std::vector<char> counted;
    long long progs_built = 0;       // leaves whose programs were (re)built

    // one HIP graph per (segment length, population-buffer parity), updated
    // to the current topology when used; a segment is the run of steps
    // between two regrids
    struct GraphEntry {
        int len, parity;
        bool judge;
        hipGraphExec_t exec;
        int topo;   // topology version the executable was built or updated for
    };
    std::vector<GraphEntry> graphs;
    int topo_version = 0;

    // an ensemble mesh (amrlbm_setup_ensemble): independent instances of two
    // roots each, every one with its own root cell size, level range, fluid
    // and bubble; ens is 0 in any other mesh
    int ens = 0;
    std::vector<amrlbm_instance> inst;
    int exec = 0;    // how a segment runs (amrlbm_set_tuning): 0 a HIP graph of two kernels per step, 1 the
                     // persistent segment kernel (when every leaf's workgroup fits on the GPU at once)
    size_t seg_lds = 0;   // the segment kernel's LDS per workgroup at the last capacity query, and
    int seg_cap = 0;      // how many such workgroups the GPU holds at once
    int order = 0;   // leaf launch order (amrlbm_set_tuning): 0 Morton in 8 XCD runs, 1 Morton, 2 ascending id,
                     // 3 as 0 with each run's heaviest ghost programs first
    Phys *d_phs = nullptr;           // [ens] device tables the kernels index by BlockDev::inst
    PhysF *d_phsf = nullptr;
    int *d_lv = nullptr;             // [ens][2] min, max levels
    double *d_iR = nullptr;          // [ens] bubble radius (initial P)
    double *d_diag_inst = nullptr;   // [ens][9]
    int ens_cap = 0;

    // the phase-field mass correction (amrlbm_set_mass_correction): at every
    // segment end the FP64 mass of each instance is put back to its budget
    // (its value when the correction was switched on plus what each mesh
    // change's restrict and prolong moved it by)
    int mass_on = 0;
    double *d_mass = nullptr;        // [instance][4]: target, S at the last fold, W, lambda
    double *d_mass_part = nullptr;   // [leaf][2] partial sums (max_blocks leaves)
    int mass_ninst = 0;

    int nlev() const { return prm.max_levels + 1; }
    int width(int L) const { return 1 << L; }
    int height(int L) const { return prm.nroots << L; }
    int tile_height(int L) const { return tile_roots << L; }
    int &cell_at(int L, int I, int J) { return at[L][(size_t)J * width(L) + I]; }
};

namespace {

// ---------------------------------------------------------------------------
// Device scratch of amrlbm_output_noalloc: the per-leaf partials (9 doubles
// per leaf, kScratchLeaves leaves at most) and the 9 totals of the FP64
// diagnostics, in a file-scope array, so that the entry point allocates
// nothing. Each GPU holds its own copy until the process ends; host code
// reaches it only through the address hipGetSymbolAddress gives on that
// device (the array's name is a host-side stand-in, not a device address).
// ---------------------------------------------------------------------------
constexpr int kScratchLeaves = 32768;
constexpr int kScratchTotals = 9 * kScratchLeaves;
static __device__ __align__(256) double scratch[kScratchTotals + 16];

struct ScratchTable {
    std::mutex mu;
    double *ptr[64] = {};   // per device id
};
ScratchTable g_scratch;

// the scratch array's address on the current device, looked up once per device
int scratch_on_current_device(double **out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(AMRLBM_ERR_HIP, "device id out of range for the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch.mu);
    if (!g_scratch.ptr[dev]) {
        double *p = nullptr;
        HIPCHK(hipGetSymbolAddress(reinterpret_cast<void **>(&p), HIP_SYMBOL(scratch)));
        g_scratch.ptr[dev] = p;
    }
    *out = g_scratch.ptr[dev];
    return AMRLBM_OK;
}

// device memory freed when its scope ends (error returns included)
struct DeviceBuffer {
    void *p = nullptr;
    ~DeviceBuffer() { (void)hipFree(p); }
    char *bytes() const { return static_cast<char *>(p); }
};

// adds the wall time of its scope (or up to stop()) to acc (ms)
struct HostTimer {
    double &acc;
    bool running = true;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    explicit HostTimer(double &a) : acc(a) {}
    void stop()
    {
        if (running) acc += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        running = false;
    }
    ~HostTimer() { stop(); }
};

// instance of a block in an ensemble mesh (its roots' pair), else 0
int inst_of(const amrlbm_mesh *m, const Block &b) { return m->ens ? (b.J >> b.level) / m->tile_roots : 0; }
// the level range that governs a block: its instance's in an ensemble mesh
int min_level_of(const amrlbm_mesh *m, const Block &b)
{
    return m->ens ? m->inst[inst_of(m, b)].min_levels : m->prm.min_levels;
}
int max_level_of(const amrlbm_mesh *m, const Block &b)
{
    return m->ens ? m->inst[inst_of(m, b)].max_levels : m->prm.max_levels;
}

// ---------------------------------------------------------------------------
// topology
// ---------------------------------------------------------------------------
int nb_value(amrlbm_mesh *m, int n, int k)   // k = 1..8
{
    static const int dsx[8] = {-1, 1, 0, 0, -1, -1, 1, 1};
    static const int dsy[8] = {0, 0, -1, 1, -1, 1, -1, 1};
    const Block &b = m->blk[n];
    int L = b.level;
    int tJ = b.J + dsy[k - 1];
    if (tJ < 0 || tJ >= m->height(L) || tJ / m->tile_height(L) != b.J / m->tile_height(L)) return -k;
    int W = m->width(L);
    int tI = ((b.I + dsx[k - 1]) % W + W) % W;
    int id = m->cell_at(L, tI, tJ);
    if (id > 0) return id;
    if (L == 0) return 0;
    return m->cell_at(L - 1, tI >> 1, tJ >> 1);   // 0 if absent: 2:1 violated
}

void place(amrlbm_mesh *m, int n)
{
    Block &b = m->blk[n];
    m->cell_at(b.level, b.I, b.J) = n;
}

void unplace(amrlbm_mesh *m, int n)
{
    Block &b = m->blk[n];
    if (b.level >= 0 && m->cell_at(b.level, b.I, b.J) == n) m->cell_at(b.level, b.I, b.J) = 0;
}

// Array_of_Leaf_Blocks (AMR_Rising_2D_OMP.f90:588-610). rows_marked: the
// caller has marked every block whose table row changed (mark_changed);
// otherwise every row goes up again.
void array_of_leaf_blocks(amrlbm_mesh *m, bool rows_marked = false)
{
    if (rows_marked && !m->rows_all) {
        // every id whose leaf status changed has its row marked: the list
        // without the marked ids, merged with those that are leaves now
        std::vector<int> &d = m->dirty_rows;
        std::sort(d.begin(), d.end());
        const std::vector<int> &old = m->leaf;
        std::vector<int> out;
        out.reserve(old.size() + d.size());
        auto it = old.begin();
        for (int id : d) {
            auto lo = std::lower_bound(it, old.end(), id);
            out.insert(out.end(), it, lo);
            it = lo;
            if (it != old.end() && *it == id) ++it;
            if (id <= m->nblks && m->blk[id].isleaf) out.push_back(id);
        }
        out.insert(out.end(), it, old.end());
        m->leaf.swap(out);
        if (std::getenv("AMRLBM_VERIFY_TABLES")) {
            std::vector<int> full;
            for (int n = 1; n <= m->nblks; n++)
                if (m->blk[n].isleaf) full.push_back(n);
            if (full != m->leaf) m->leaf_list_bad = true;
        }
    } else {
        m->leaf.clear();
        for (int n = 1; n <= m->nblks; n++)
            if (m->blk[n].isleaf) m->leaf.push_back(n);
    }
    m->tables_dirty = true;
    if (!rows_marked) {   // a topology from elsewhere: its flags as given
        m->rows_all = true;
        m->flagged.clear();
        for (int n = 1; n <= m->nblks; n++)
            if (m->blk[n].refin || m->blk[n].deref) m->flagged.push_back(n);
    }
    m->topo_version++;
}

// block id's table row to go up with the next upload
void mark_row(amrlbm_mesh *m, int id)
{
    if (id < 1) return;
    if ((int)m->dirty_mark.size() <= id) m->dirty_mark.resize((size_t)id + 1025, 0);
    if (!m->dirty_mark[id]) {
        m->dirty_mark[id] = 1;
        m->dirty_rows.push_back(id);
    }
}

// The table rows a change of block t's leaf status or children can touch: its
// own, its parent's (hasLeaf) and those of every block whose neighbour entries
// may name t or its children: the blocks nb_value gives t (its level or
// coarser), their parents and their children. Under 2:1 balance no finer block
// touches t. Called on the final topology of a regrid, for every parent refined
// or derefined and every child created; a freed child's row is marked alone.
void mark_changed(amrlbm_mesh *m, int t)
{
    mark_row(m, t);
    const Block &b = m->blk[t];
    if (b.level < 0) return;
    if (b.parent > 0) mark_row(m, b.parent);
    for (int k = 1; k <= 8; k++) {
        const int v = nb_value(m, t, k);
        if (v <= 0) continue;
        mark_row(m, v);
        const Block &q = m->blk[v];
        if (q.parent > 0) mark_row(m, q.parent);
        if (!q.isleaf && q.has_child)
            for (int c = 0; c < 4; c++) mark_row(m, q.child[c]);
    }
}

// Determine_Min_Max_Levels (AMR_Rising_2D_OMP.f90:434-452)
void determine_min_max_levels(amrlbm_mesh *m)
{
    int lmin = 100, lmax = -100;
    for (int n : m->leaf) {
        lmin = std::min(lmin, m->blk[n].level);
        lmax = std::max(lmax, m->blk[n].level);
    }
    m->levelmin = lmin;
    m->levelmax = lmax;
}

// Add_Child_Blocks (AMR_Rising_2D_OMP.f90:517-584); the neighbour pointers of
// Find_Neighbors are implied by the position map
int add_child_blocks(amrlbm_mesh *m, int par, int freeblk)
{
    int N = freeblk - 1;
    if (N + 4 > m->prm.max_blocks)
        return fail(AMRLBM_ERR_CAPACITY, "number of blocks exceeded MaxBlks (" + std::to_string(m->prm.max_blocks) + ")");
    Block &P = m->blk[par];
    for (int i = 0; i < 4; i++) P.child[i] = N + 1 + i;
    P.isleaf = false;
    P.hasleaf = true;
    if (P.parent > 0) m->blk[P.parent].hasleaf = false;
    for (int J = 0; J <= 1; J++)
        for (int I = 0; I <= 1; I++) {
            N++;
            Block &b = m->blk[N];
            b.level = P.level + 1;
            b.dx = P.dx / 2;
            b.I = 2 * P.I + I;
            b.J = 2 * P.J + J;
            b.parent = par;
            b.has_child = true;
            for (int k = 0; k < 4; k++) b.child[k] = N;
            b.isleaf = true;
            b.hasleaf = false;
            b.refin = false;
            b.deref = false;
            b.alloc = true;
            place(m, N);
        }
    m->nblks = std::max(m->nblks, N);
    return AMRLBM_OK;
}

// Remove_Children (Refine_Derefine.f90:215-258)
void remove_children(amrlbm_mesh *m, int par)
{
    int c1 = m->blk[par].child[0];
    for (int c = c1; c <= c1 + 3; c++) {
        unplace(m, c);
        m->blk[c].level = -1;
        m->blk[c].isleaf = false;
        m->blk[c].has_child = false;
        m->blk[c].alloc = false;
    }
    Block &P = m->blk[par];
    for (int k = 0; k < 4; k++) P.child[k] = par;
    P.isleaf = true;
    P.hasleaf = false;
    int gp = P.parent;
    if (gp > 0) {
        bool all = true;
        for (int s = m->blk[gp].child[0]; s <= m->blk[gp].child[3]; s++) all = all && m->blk[s].isleaf;
        if (all) m->blk[gp].hasleaf = true;
    }
    m->freeq.push_back(c1);
    m->kk += 4;
}

// ---------------------------------------------------------------------------
// ghost programs
// ---------------------------------------------------------------------------
// One axis of Biquadratic_from_Fine as the directional fills use it: source
// cells in the 32-wide fine frame of the refined neighbour, listed from the
// interface outward (normal axis) or upward (tangential axis); weight pattern
// 0 = (3,6,-1), 1 = (-1,6,3) by the parity of the coarse index
// (Interpolations.f90:13-31; Fill_Left_hg, Fill_Ghost_Cells_hg.f90:74-85).
int fine_axis(int s, int t, int f[3])
{
    int k = s == 0 ? t : (s < 0 ? 1 - t : t - NX);
    int m0 = (k & 1) ? 2 * k - 1 : 2 * k - 2;
    for (int a = 0; a < 3; a++) f[a] = s < 0 ? 2 * NX + 1 - (m0 + a) : m0 + a;
    return (k & 1) ? 0 : 1;
}

// One axis of Biquadratic_from_Coarse: coarse cells around the fine cell,
// the first lying toward the block being filled on a normal axis
// (Fill_Left_hg:91-96, Fill_Left_Bottom_hg:311-315); weight pattern
// 2 = (5,30,-3), 3 = (-3,30,5), the 5 on the side the fine cell leans to
// (Interpolations.f90:85-101). pos = the block's position in its parent.
int coarse_axis(int s, int t, int pos, int c[3])
{
    int fp = pos * NX + t;
    int cp = (fp + 1) >> 1;
    int sn = cp < 1 ? -1 : (cp > NX ? 1 : 0);
    int cc = cp - sn * NX;
    bool upper = (fp & 1) == 0;
    if (s < 0) {
        c[0] = cc + 1; c[1] = cc; c[2] = cc - 1;
        return upper ? 2 : 3;
    }
    c[0] = cc - 1; c[1] = cc; c[2] = cc + 1;
    return upper ? 3 : 2;
}

int nb_value(amrlbm_mesh *m, int n, int k);

// Builds the ghost program of leaf n for one fill kind. Every ghost cell the
// reference fills for n (Fill_Ghost_Cells_{C,C2,hg}: the eight directional
// fills, then the wall rows by BC_Cell.f90) becomes one node; a case (iii)
// fill reads the coarser neighbour's ghost cells, which in the reference were
// filled in the coarser level's pass, so those cells become nodes too,
// recursively, in the same operation order. Interior values, the block's own
// included, are read from memory (left by the previous kernel), everything
// else from LDS; so the first stage reads memory only and a kernel can
// evaluate it before its first barrier.
struct ProgBuild {
    amrlbm_mesh *m;
    int kind, R, n;
    // the node list, each node's stage and the stage of the node that writes
    // LDS cell c: the host thread's, reused by every build on that thread
    struct Lists {
        std::vector<PNode> nodes;
        std::vector<int> nstage;
        std::vector<int> cell_stage = std::vector<int>((FS > CS ? FS : CS) + NAUX_MAX, 0);
    };
    static Lists &fresh_lists()
    {
        thread_local Lists tl;
        tl.nodes.clear();
        tl.nstage.clear();
        return tl;
    }
    Lists &lists = fresh_lists();
    std::vector<PNode> &nodes = lists.nodes;
    std::vector<int> &nstage = lists.nstage;
    std::vector<int> &cell_stage = lists.cell_stage;   // sized for every dest; read only once written
    // the neighbours (nb_value) of the blocks this build reads, asked for up to
    // three times per node: direct-mapped by id, each k filled when first asked
    static constexpr int NBC = 32, NB_UNSET = 1 << 30;
    int nbc_id[NBC] = {};
    int nbc_v[NBC][8];
    int nbv(int b, int k)
    {
        const int h = b & (NBC - 1);
        if (nbc_id[h] != b) {
            nbc_id[h] = b;
            for (int j = 0; j < 8; j++) nbc_v[h][j] = NB_UNSET;
        }
        int &v = nbc_v[h][k - 1];
        if (v == NB_UNSET) v = nb_value(m, b, k);
        return v;
    }
    // node memo: open addressing on (block, x, y); -2 marks a node being built.
    // The table is the host thread's, reused by every build on that thread: a
    // slot belongs to this build iff its stamp is the build's.
    static constexpr int MEMO = 2048;
    struct Memo {
        std::vector<long long> key = std::vector<long long>(MEMO, -1);
        std::vector<int> val = std::vector<int>(MEMO, -1);
        std::vector<unsigned> stamp = std::vector<unsigned>(MEMO, 0);
        unsigned now = 0;
    };
    Memo &memo = fresh_memo();
    static Memo &fresh_memo()
    {
        thread_local Memo tm;
        if (++tm.now == 0) {   // stamp wrap: clear once
            std::fill(tm.stamp.begin(), tm.stamp.end(), 0u);
            tm.now = 1;
        }
        return tm;
    }
    int naux = 0;
    std::string err;
    // every block the build reads (see ProgCache), each once: a block goes
    // into deps when its mark is not the caller's stamp yet
    std::vector<int> *deps = nullptr;
    unsigned *dep_mark = nullptr;
    unsigned dep_stamp = 0;
    void dep(int b)
    {
        if (deps && b > 0 && dep_mark[b] != dep_stamp) {
            dep_mark[b] = dep_stamp;
            deps->push_back(b);
        }
    }

    int own_cells() const { return kind == KIND_HG ? FS : CS; }
    int cell_index(int x, int y) const { return kind == KIND_HG ? y * FW + x : (y + 1) * CW + (x + 1); }
    static bool interior(int x, int y) { return x >= 1 && x <= NX && y >= 1 && y <= NY; }
    int global_ref(int b, int x, int y) const
    {
        return kind == KIND_HG ? (b - 1) * POPB + pcell(x, y) : (b - 1) * CS + cell_index(x, y);
    }
    bool memo_used(int h) const { return memo.stamp[h] == memo.now; }
    int memo_slot(long long key) const
    {
        unsigned h = (unsigned)((key * 0x9E3779B97F4A7C15ull) >> 40) & (MEMO - 1);
        while (memo_used(h) && memo.key[h] != key) h = (h + 1) & (MEMO - 1);
        return (int)h;
    }
    int ref(int b, int x, int y, int &st)
    {
        st = 0;
        dep(b);
        if (interior(x, y)) return global_ref(b, x, y);
        int dest = node(b, x, y);
        if (dest < 0) return 0;
        st = cell_stage[dest];
        return -(1 + dest);
    }
    int node(int b, int x, int y)
    {
        const long long key = ((long long)b << 16) | ((long long)(x + 8) << 8) | (long long)(y + 8);
        const int slot = memo_slot(key);
        if (memo_used(slot)) {
            if (memo.val[slot] == -2) {
                err = "cyclic ghost dependency";
                return -1;
            }
            return memo.val[slot];
        }
        memo.stamp[slot] = memo.now;
        memo.key[slot] = key;
        memo.val[slot] = -2;
        PNode nd{};
        int st = 0, s2 = 0;
        dep(b);
        const Block &B = m->blk[b];
        if (y <= 0) dep(nbv(b, 3));
        if (y >= NY + 1) dep(nbv(b, 4));
        const bool bot = y <= 0 && nbv(b, 3) < 0;
        const bool top = y >= NY + 1 && nbv(b, 4) < 0;
        if (bot || top) {
            if (kind == KIND_HG) {   // Bounce_Back_3/4 (BC_Cell.f90:2-42)
                const int yi = bot ? 1 : NY;
                nd.info = bot ? NT_BB_BOT : NT_BB_TOP;
                for (int a = 0; a < 3; a++) {
                    int xn = std::min(std::max(x - 1 + a, 0), NX + 1);
                    nd.ref[a] = ref(b, xn, yi, s2);
                    st = std::max(st, s2);
                }
            } else {                 // Symmetric_C_3/4, Symmetric_C2_3/4 (BC_Cell.f90:46-104)
                const int ys = bot ? 1 - y : 2 * NY + 1 - y;
                nd.info = NT_COPY;
                nd.ref[0] = ref(b, x, ys, st);
            }
        } else {
            static const int dtab[3][3] = {{4, 0, 5}, {2, -1, 3}, {6, 1, 7}};   // [sx+1][sy+1]
            const int sx = x < 1 ? -1 : (x > NX ? 1 : 0), sy = y < 1 ? -1 : (y > NY ? 1 : 0);
            const int d = dtab[sx + 1][sy + 1];
            const int N1 = d < 0 ? 0 : nbv(b, d + 1);
            dep(N1);
            if (d < 0 || N1 <= 0) {
                err = "ghost cell (" + std::to_string(x) + "," + std::to_string(y) + ") of block " + std::to_string(b) +
                      " has no neighbour to fill it";
                return -1;
            }
            const Block &Q = m->blk[N1];
            if (Q.level == B.level && Q.isleaf) {   // case (i): copy (Fill_Left_hg:66-70)
                nd.info = NT_COPY;
                nd.ref[0] = ref(N1, x - sx * NX, y - sy * NY, st);
            } else if (Q.level == B.level) {        // case (ii): from the children (Fill_Left_hg:72-85)
                int fx[3], fy[3];
                int px = fine_axis(sx, x, fx), py = fine_axis(sy, y, fy);
                nd.info = NT_FINE | (px << 4) | (py << 6);
                for (int a = 0; a < 3; a++) {
                    int ix = fx[a] > NX, lx = fx[a] - ix * NX;
                    for (int c = 0; c < 3; c++) {
                        int iy = fy[c] > NY, ly = fy[c] - iy * NY;
                        nd.ref[3 * a + c] = ref(Q.child[0] + ix + 2 * iy, lx, ly, s2);
                        st = std::max(st, s2);
                    }
                }
            } else if (B.level == Q.level + 1) {    // case (iii): from the coarser block (Fill_Left_hg:89-99)
                int cx[3], cy[3];
                int px = coarse_axis(sx, x, B.I & 1, cx), py = coarse_axis(sy, y, B.J & 1, cy);
                nd.info = NT_COARSE | (px << 4) | (py << 6);
                for (int a = 0; a < 3; a++)
                    for (int c = 0; c < 3; c++) {
                        nd.ref[3 * a + c] = ref(N1, cx[a], cy[c], s2);
                        st = std::max(st, s2);
                    }
            } else {
                err = "level jump between blocks " + std::to_string(b) + " and " + std::to_string(N1);
                return -1;
            }
        }
        if (!err.empty()) return -1;
        int dest;
        if (b == n) {
            dest = cell_index(x, y);
        } else {
            if (naux == NAUX_MAX) {
                err = "more than " + std::to_string(NAUX_MAX) + " auxiliary ghost cells";
                return -1;
            }
            dest = own_cells() + naux++;
        }
        nd.dest = dest;
        nd.own = (kind == KIND_HG && b == n) ? pcell(x, y) : -1;
        cell_stage[dest] = st + 1;
        nodes.push_back(nd);
        nstage.push_back(st + 1);
        memo.val[slot] = dest;
        return dest;
    }
    int build()
    {
        for (int y = 1 - R; y <= NY + R; y++)
            for (int x = 1 - R; x <= NX + R; x++)
                if (!interior(x, y) && node(n, x, y) < 0) return -1;
        return 0;
    }
};


// deps (optional): the blocks the program reads are appended, each once per
// value of stamp (dep_mark holds a stamp per block id)
void build_leaf_program(amrlbm_mesh *m, int kind, int n, LeafProg &out, std::vector<int> *deps = nullptr,
                        unsigned *dep_mark = nullptr, unsigned stamp = 0)
{
    out.nodes.clear();
    out.err.clear();
    out.naux = out.nst = 0;
    std::fill(out.end, out.end + MAX_STAGES, 0);
    ProgBuild pb;
    pb.m = m;
    pb.kind = kind;
    pb.R = kind == KIND_C2 ? 2 : 1;
    pb.n = n;
    pb.deps = deps;
    pb.dep_mark = dep_mark;
    pb.dep_stamp = stamp;
    pb.dep(n);
    if (pb.build()) {
        out.err = "ghost program of block " + std::to_string(n) + ": " + pb.err;
        return;
    }
    // every reference must land on an interior cell of a live block or on a staged LDS cell
    const int slotsz = kind == KIND_HG ? POPB : CS, own = pb.own_cells();
    for (const PNode &nd : pb.nodes) {
        const int type = nd.info & 15;
        const int nref = type == NT_COPY ? 1 : (type >= NT_BB_BOT ? 3 : 9);
        bool ok = nd.dest >= 0 && nd.dest < own + NAUX_MAX;
        for (int r = 0; r < nref && ok; r++) {
            int v = nd.ref[r];
            if (v < 0) {
                ok = -1 - v < own + pb.naux;
            } else {
                const int q = kind == KIND_HG ? v / POPB : v / CS;   // constant divisors
                int b = q + 1, c = v - q * slotsz;
                bool in = kind == KIND_HG ? c < NX * NY : ProgBuild::interior(c % CW - 1, c / CW - 1);
                ok = b <= m->nblks && m->blk[b].level >= 0 && m->blk[b].isleaf && in;
            }
        }
        if (!ok) {
            out.err = "ghost program of block " + std::to_string(n) + " references a cell outside the mesh";
            return;
        }
    }
    int nst = 0;
    for (int v : pb.nstage) nst = std::max(nst, v);
    if (nst > MAX_STAGES) {
        out.err = "ghost program of block " + std::to_string(n) + " deeper than MAX_STAGES";
        return;
    }
    const size_t cap = kind == KIND_HG ? NODES_HG : (kind == KIND_C2 ? NODES_C2 : NODES_C1);
    if (pb.nodes.size() > cap) {
        out.err = "ghost program of block " + std::to_string(n) + " has more recipes than a kernel stages";
        return;
    }
    out.nst = nst;
    out.naux = pb.naux;
    // stage by stage, each in build order
    int cnt[MAX_STAGES + 1] = {};
    for (int v : pb.nstage) cnt[v]++;
    int pos[MAX_STAGES + 1] = {};
    for (int st = 1; st <= nst; st++) {
        pos[st] = st == 1 ? 0 : out.end[st - 2];
        out.end[st - 1] = pos[st] + cnt[st];
    }
    out.nodes.resize(pb.nodes.size());
    for (size_t i = 0; i < pb.nodes.size(); i++) out.nodes[pos[pb.nstage[i]]++] = pb.nodes[i];
}

// nodes per block slot in the device arenas, by kind (C1, C2, HG): the most
// a program may hold (build_leaf_program refuses more)
constexpr int kProgSlot[3] = {NODES_C1, NODES_C2, NODES_HG};

bool cache_current(const amrlbm_mesh *m, int id)
{
    const ProgCache &pc = m->pcache[id];
    if (!pc.valid) return false;
    for (const auto &d : pc.deps)
        if (!(block_sig(m->blk[d.first]) == d.second)) return false;
    return true;
}

// Programs of every leaf for the three kinds. A leaf's programs are rebuilt
// only when a block they were built from changed (ProgCache); the builds only
// read the topology, so they run on several host threads. With
// AMRLBM_VERIFY_PROGRAMS set in the environment every reused program is
// checked against a fresh build.
// the host threads for a pass over nl leaves: one below 4096 leaves, where
// waking a team costs more than the pass
int prog_threads(int nl) { return nl < 4096 ? 1 : std::min(omp_get_max_threads(), nl / 512); }

// operand loads of a program: the cost of a block's ghost fill
int program_operands(const LeafProg &lp)
{
    int ops = 0;
    for (const PNode &nd : lp.nodes) {
        const int type = nd.info & 15;
        ops += type == NT_COPY ? 1 : (type >= NT_BB_BOT ? 3 : 9);
    }
    return ops;
}

// one id's aux-cell and recipe counts into (sign 1) or out of (-1) the
// per-kind histograms over the leaves
void count_programs(amrlbm_mesh *m, int id, int sign)
{
    for (int kind = 0; kind < 3; kind++) {
        m->hist_aux[kind][m->naux_id[kind][id]] += sign;
        m->hist_nn[kind][m->nn_id[kind][id]] += sign;
    }
    m->counted[id] = sign > 0;
}

int rebuild_programs(amrlbm_mesh *m)
{
    HostTimer timer(m->host_ms_programs);
    const int nl = (int)m->leaf.size();
    if (m->pcache.size() < m->blk.size()) m->pcache.resize(m->blk.size());
    const int nthr = prog_threads(nl);
    // one team for both passes: the check of every leaf's cache (about 0.03 us
    // per leaf), then the builds of the stale ones (about 20 us per leaf, one
    // leaf at a time so that a few dirty leaves spread over the team). Eight
    // threads: 20,000 steps of the 1024x2048 rising bubble took 523 ms with 8, 546 ms with 16
    // (the idle team spins between regrids and takes host time from the step
    // loop) and 568 ms with 1.
    const int team = nl >= 128 ? std::min(omp_get_max_threads(), 8) : 1;
    if (m->rdeps.size() < m->blk.size()) m->rdeps.resize(m->blk.size());
    // the leaves to check: every one after a full upload, else those that read
    // a block whose row changed, and the changed blocks that are leaves now
    std::vector<int> cand;
    if (m->check_all) {
        cand = m->leaf;
    } else {
        for (int b : m->changed_blocks) {
            if (m->blk[b].isleaf) cand.push_back(b);
            for (int id : m->rdeps[b])
                if (m->blk[id].isleaf) cand.push_back(id);
        }
        std::sort(cand.begin(), cand.end());
        cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    }
    m->changed_blocks.clear();
    m->check_all = false;
    const int nc = (int)cand.size();
    std::vector<char> stale(nc);
    std::vector<int> dirty;
    std::vector<std::vector<std::pair<int, BlockSig>>> old_deps;
    int nd = 0;
#pragma omp parallel num_threads(team)
    {
#pragma omp for schedule(static)
        for (int k = 0; k < nc; k++) stale[k] = !cache_current(m, cand[k]);
#pragma omp single
        {
            for (int k = 0; k < nc; k++)
                if (stale[k]) dirty.push_back(cand[k]);
            nd = (int)dirty.size();
            old_deps.resize(nd);
            for (int i = 0; i < nd; i++) old_deps[i] = std::move(m->pcache[dirty[i]].deps);
        }
        std::vector<int> d;
        // per block id, the stamp of the last leaf whose build listed it
        thread_local std::vector<unsigned> mark;
        thread_local unsigned stamp = 0;
        if (mark.size() < m->blk.size()) mark.resize(m->blk.size(), 0u);
#pragma omp for schedule(dynamic, 1)
        for (int i = 0; i < nd; i++) {   // one leaf's three programs, then the blocks they read
            ProgCache &pc = m->pcache[dirty[i]];
            pc.valid = false;
            pc.on_device = false;
            d.clear();
            if (++stamp == 0) {   // wrap: clear once
                std::fill(mark.begin(), mark.end(), 0u);
                stamp = 1;
            }
            for (int kind = 0; kind < 3; kind++) build_leaf_program(m, kind, dirty[i], pc.p[kind], &d, mark.data(), stamp);
            std::sort(d.begin(), d.end());
            pc.deps.clear();
            for (int b : d) pc.deps.push_back({b, block_sig(m->blk[b])});
        }
    }
    for (int kind = 0; kind < 3; kind++)
        if (m->hdr_id[kind].size() < m->pcache.size()) {
            m->hdr_id[kind].resize(m->pcache.size());
            m->naux_id[kind].resize(m->pcache.size(), 0);
            m->nn_id[kind].resize(m->pcache.size(), 0);
        }
    if (m->cost_id.size() < m->pcache.size()) m->cost_id.resize(m->pcache.size(), 0);
    if (m->counted.size() < m->pcache.size()) m->counted.resize(m->pcache.size(), 0);
    for (int i = 0; i < nd; i++) {
        const int id = dirty[i];
        ProgCache &pc = m->pcache[id];
        if (m->counted[id]) count_programs(m, id, -1);   // its counts change below
        for (int kind = 0; kind < 3; kind++) {
            const LeafProg &lp = pc.p[kind];
            if (!lp.err.empty()) return fail(AMRLBM_ERR_TOPOLOGY, lp.err);
            PHdr &h = m->hdr_id[kind][id];
            h.off = (id - 1) * kProgSlot[kind];
            h.nst = lp.nst;
            for (int st = 0; st < MAX_STAGES; st++) h.end[st] = lp.end[st];
            m->naux_id[kind][id] = lp.naux;
            m->nn_id[kind][id] = (int)lp.nodes.size();
        }
        m->cost_id[id] = program_operands(pc.p[KIND_C2]) + program_operands(pc.p[KIND_HG]);
        pc.valid = true;
    }
    m->prog_pending.insert(m->prog_pending.end(), dirty.begin(), dirty.end());
    // the reverse index follows each rebuilt program's dependencies
    for (int i = 0; i < nd; i++) {
        const int id = dirty[i];
        for (const auto &dep : old_deps[i]) {
            std::vector<int> &r = m->rdeps[dep.first];
            auto it = std::find(r.begin(), r.end(), id);
            if (it != r.end()) {
                *it = r.back();
                r.pop_back();
            }
        }
        for (const auto &dep : m->pcache[id].deps) m->rdeps[dep.first].push_back(id);
    }
    m->progs_built += nd;
    if (std::getenv("AMRLBM_VERIFY_PROGRAMS")) {
        std::vector<LeafProg> fresh((size_t)3 * nl);
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthr)
        for (int t = 0; t < 3 * nl; t++) build_leaf_program(m, t % 3, m->leaf[t / 3], fresh[t]);
        for (int t = 0; t < 3 * nl; t++) {
            const LeafProg &a = fresh[t], &b = m->pcache[m->leaf[t / 3]].p[t % 3];
            bool same = a.err.empty() && a.nst == b.nst && a.naux == b.naux && a.nodes.size() == b.nodes.size() &&
                        std::equal(a.end, a.end + MAX_STAGES, b.end) &&
                        std::memcmp(a.nodes.data(), b.nodes.data(), sizeof(PNode) * a.nodes.size()) == 0;
            if (!same)
                return fail(AMRLBM_ERR_TOPOLOGY, "reused ghost program of block " + std::to_string(m->leaf[t / 3]) +
                                                     " differs from a fresh build");
        }
    }
    return AMRLBM_OK;
}

// The programs of every leaf on the device: the rebuilt ones into their
// slots, the headers in launch order (m->disp).
// The pinned batch buffer is written only once the previous batch's scatter
// (and the restrict / prolong launches queued before it) has run.
int batch_begin(amrlbm_mesh *m)
{
    if (m->tab_recorded) HIPCHK(hipEventSynchronize(m->ev_tab));
    m->up_used = 0;
    m->segs.clear();
    return AMRLBM_OK;
}

// bytes of batch space, 16-byte aligned; *at = its offset. A pointer into
// the buffer is valid until the next batch_alloc (the buffer may grow).
int batch_alloc(amrlbm_mesh *m, size_t bytes, size_t *at)
{
    const size_t off = (m->up_used + 15) & ~(size_t)15;
    if (off + bytes > m->up_cap) {
        const size_t cap = (off + bytes) + (off + bytes) / 2 + 65536;
        char *p = nullptr;
        HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&p), cap, hipHostMallocDefault));
        if (m->up_used) std::memcpy(p, m->h_up, m->up_used);
        if (m->h_up) HIPCHK(hipHostFree(m->h_up));
        m->h_up = p;
        m->up_cap = cap;
    }
    m->up_used = off + bytes;
    *at = off;
    return AMRLBM_OK;
}

// batch bytes [at, at + bytes) go to device address dst, in segments of at
// most 8 KB (one workgroup each), so that a large table spreads over the GPU
void batch_seg(amrlbm_mesh *m, void *dst, size_t at, size_t bytes)
{
    constexpr size_t kSeg = 8192;
    char *d = static_cast<char *>(dst);
    for (size_t o = 0; o < bytes; o += kSeg) {
        const size_t n = std::min(kSeg, bytes - o);
        m->segs.push_back({(unsigned long long)(uintptr_t)(d + o), (int)((at + o) / 4), (int)(n / 4)});
    }
}

// one copy of the batch and one scatter launch; no wait (batch_begin waits)
int batch_flush(amrlbm_mesh *m)
{
    if (m->segs.empty()) return AMRLBM_OK;
    size_t at = 0;
    const size_t ns = m->segs.size();
    int rc = batch_alloc(m, sizeof(ScatterSeg) * ns, &at);
    if (rc) return rc;
    std::memcpy(m->h_up + at, m->segs.data(), sizeof(ScatterSeg) * ns);
    if (m->up_used > m->dup_cap) {
        if (m->d_up) HIPCHK(hipFree(m->d_up));
        m->d_up = nullptr;
        m->dup_cap = m->up_used + m->up_used / 2 + 65536;
        HIPCHK(hipMalloc(&m->d_up, m->dup_cap));
    }
    HIPCHK(hipMemcpyAsync(m->d_up, m->h_up, m->up_used, hipMemcpyHostToDevice, m->stream));
    launch_scatter_segments(m->d_up, reinterpret_cast<const ScatterSeg *>(static_cast<char *>(m->d_up) + at), (int)ns,
                            m->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(m->ev_tab, m->stream));
    m->tab_recorded = true;
    return AMRLBM_OK;
}

int upload_programs(amrlbm_mesh *m)
{
    HostTimer timer(m->host_ms_programs);
    const int nl = (int)m->leaf.size();
    // the arenas hold a slot per block id up to nblks; grown (and refilled) as ids pass it
    if (m->nblks > m->prog_slots) {
        const int cap = std::min(m->prm.max_blocks, std::max(m->nblks + m->nblks / 2, 1024));
        for (int kind = 0; kind < 3; kind++) {
            if (m->d_pnode[kind]) HIPCHK(hipFree(m->d_pnode[kind]));
            m->d_pnode[kind] = nullptr;
            HIPCHK(hipMalloc(&m->d_pnode[kind], sizeof(PNode) * kProgSlot[kind] * (size_t)cap));
        }
        m->prog_slots = cap;
        for (ProgCache &pc : m->pcache) pc.on_device = false;
        m->leaf_delta_all = true;
    }
    // the leaves whose programs are not in their slots yet, ascending id: after
    // a full upload or a regrown arena any leaf, else a rebuilt one or one that
    // became a leaf (its row changed) with its cached programs; the aux-cell and
    // recipe histograms follow the same ids
    const bool all = m->leaf_delta_all || m->hist_aux[0].empty();
    std::vector<int> up;
    if (all) {
        for (int n : m->leaf)
            if (!m->pcache[n].on_device) up.push_back(n);
        for (int kind = 0; kind < 3; kind++) {
            m->hist_aux[kind].assign(NAUX_MAX + 1, 0);
            m->hist_nn[kind].assign(kProgSlot[kind] + 1, 0);
        }
        std::fill(m->counted.begin(), m->counted.end(), 0);
        for (int n : m->leaf) count_programs(m, n, 1);
    } else {
        std::vector<int> &c = m->prog_pending;
        c.insert(c.end(), m->leaf_delta.begin(), m->leaf_delta.end());
        std::sort(c.begin(), c.end());
        c.erase(std::unique(c.begin(), c.end()), c.end());
        for (int n : c) {
            const bool lf = n <= m->nblks && m->blk[n].isleaf;
            if (lf && !m->pcache[n].on_device) up.push_back(n);
            if (m->counted[n] && !lf) count_programs(m, n, -1);
            else if (!m->counted[n] && lf) count_programs(m, n, 1);
        }
    }
    m->prog_pending.clear();
    m->leaf_delta_all = false;
    if (std::getenv("AMRLBM_VERIFY_PROGRAMS")) {   // the kept lists against passes over every leaf
        for (int n : m->leaf)
            if (!m->pcache[n].on_device && !std::binary_search(up.begin(), up.end(), n))
                return fail(AMRLBM_ERR_TOPOLOGY, "leaf " + std::to_string(n) + " has no programs on the device");
        for (int kind = 0; kind < 3; kind++) {
            int naux = 0, nmax = 0;
            for (int id : m->leaf) {
                naux = std::max(naux, m->naux_id[kind][id]);
                nmax = std::max(nmax, m->nn_id[kind][id]);
            }
            int ha = NAUX_MAX, hn = kProgSlot[kind];
            while (ha > 0 && !m->hist_aux[kind][ha]) ha--;
            while (hn > 0 && !m->hist_nn[kind][hn]) hn--;
            if (ha != naux || hn != nmax)
                return fail(AMRLBM_ERR_TOPOLOGY, "kept program counts differ from a pass over the leaves");
        }
    }
    const int nup = (int)up.size();
    for (int kind = 0; kind < 3; kind++) {
        // the LDS the kernels size by: the largest counts present
        int naux = NAUX_MAX, nmax = kProgSlot[kind];
        while (naux > 0 && !m->hist_aux[kind][naux]) naux--;
        while (nmax > 0 && !m->hist_nn[kind][nmax]) nmax--;
        if ((size_t)nl > m->phdr_cap[kind]) {
            if (m->d_phdr[kind]) HIPCHK(hipFree(m->d_phdr[kind]));
            m->d_phdr[kind] = nullptr;
            m->phdr_cap[kind] = nl + nl / 4 + 64;
            HIPCHK(hipMalloc(&m->d_phdr[kind], sizeof(PHdr) * m->phdr_cap[kind]));
        }
        // the nodes of the leaves in up, packed, each a segment to its block slot
        std::vector<size_t> at(nup + 1, 0);
        for (int i = 0; i < nup; i++) at[i + 1] = at[i] + m->pcache[up[i]].p[kind].nodes.size();
        const size_t total = at[nup];
        size_t base = 0;
        int rc = batch_alloc(m, sizeof(PNode) * total, &base);
        if (rc) return rc;
        PNode *nodes = reinterpret_cast<PNode *>(m->h_up + base);
        // a few hundred KB after a regrid: copied by the team the builds just used
#pragma omp parallel for schedule(dynamic, 1) num_threads(nup >= 8 ? std::min(omp_get_max_threads(), 8) : 1)
        for (int i = 0; i < nup; i++) {
            const LeafProg &lp = m->pcache[up[i]].p[kind];
            std::memcpy(nodes + at[i], lp.nodes.data(), sizeof(PNode) * lp.nodes.size());
        }
        for (int i = 0; i < nup; i++)
            batch_seg(m, m->d_pnode[kind] + (size_t)(up[i] - 1) * kProgSlot[kind], base + sizeof(PNode) * at[i],
                      sizeof(PNode) * (at[i + 1] - at[i]));
        // the headers of the leaves in up, each to its block slot; the kernels'
        // launch-order headers are gathered from them on the device
        // (launch_gather_headers, after the batch)
        rc = batch_alloc(m, sizeof(PHdr) * nup, &base);
        if (rc) return rc;
        PHdr *hdr = reinterpret_cast<PHdr *>(m->h_up + base);
        for (int i = 0; i < nup; i++) hdr[i] = m->hdr_id[kind][up[i]];
        for (int i = 0; i < nup;) {   // runs of consecutive ids, one segment each
            int j = i + 1;
            while (j < nup && up[j] == up[j - 1] + 1) j++;
            batch_seg(m, m->d_phdr_slot[kind] + (up[i] - 1), base + sizeof(PHdr) * i, sizeof(PHdr) * (j - i));
            i = j;
        }
        m->prog[kind].hdr = m->d_phdr[kind];
        m->prog[kind].nodes = m->d_pnode[kind];
        m->prog[kind].naux = naux;   // the hg kernel sizes its LDS by these
        m->prog[kind].nmax = nmax;
        m->prog_nodes_uploaded += (long long)total;
    }
    for (int id : up) m->pcache[id].on_device = true;   // in stream order: the batch scatter precedes any use
    return AMRLBM_OK;
}

// ---------------------------------------------------------------------------
// device tables
// ---------------------------------------------------------------------------
int ensure_int_buffer(int **buf, int *cap, int need)
{
    if (need <= *cap) return AMRLBM_OK;
    if (*buf) HIPCHK(hipFree(*buf));
    int c = std::max(need, 1024);
    HIPCHK(hipMalloc(buf, sizeof(int) * c));
    *cap = c;
    return AMRLBM_OK;
}

// Launch order of the leaves. A launch hands workgroup k to XCD (k + s) % 8,
// s the XCD its first workgroup lands on (it varies from launch to launch),
// and each XCD has its own L2, so the leaves go along a Morton curve
// (finest-cell origin of the block), the curve is cut into 8 runs of
// neighbouring blocks, and run c takes the positions k = c mod 8: a block and
// most of its neighbours then run on one XCD in every launch. In an ensemble mesh the
// curve runs instance by instance, each in its own frame (instances differ
// in cell size, so their absolute origins overlap).
// a leaf's Morton key: its instance above bit 40, then the finest-cell origin
// of the block in the instance's frame, x and y bits interleaved
unsigned long long morton_key(const amrlbm_mesh *m, int id)
{
    const Block &b = m->blk[id];
    const int q = inst_of(m, b), J = b.J - q * (m->tile_roots << b.level);
    const unsigned x = (unsigned)(b.I * b.dx * NX), y = (unsigned)(J * b.dx * NY);
    // an ensemble's frames lie below 2^19 (L0 <= nx 2^14, two roots), leaving bits 40.. for the instance
    const int nbits = m->ens ? 20 : 31;
    unsigned long long z = 0;
    for (int bit = 0; bit < nbits; bit++)
        z |= (unsigned long long)((x >> bit) & 1u) << (2 * bit) | (unsigned long long)((y >> bit) & 1u) << (2 * bit + 1);
    return (unsigned long long)q << 40 | z;
}

void dispatch_order(amrlbm_mesh *m)
{
    const int n = (int)m->leaf.size();
    if (m->order == 2) {   // ascending id (Array_of_Leaf_Blocks); the sorted keys are not kept meanwhile
        m->disp = m->leaf;
        m->morton_all = true;
        m->morton_changed.clear();
        return;
    }
    // the sorted keys: after a mesh change, the entries of the blocks whose rows
    // changed are dropped, and those that are leaves now come back re-keyed, in
    // one merge pass
    std::vector<std::pair<unsigned long long, int>> &sorted = m->morton;
    if (!m->morton_all) {
        std::vector<char> &drop = m->morton_drop;   // all zero between calls
        if (drop.size() < (size_t)m->nblks + 1) drop.resize((size_t)m->nblks + 1025, 0);
        std::vector<std::pair<unsigned long long, int>> add;
        std::vector<int> dropped;
        for (int id : m->morton_changed) {
            if (id > m->nblks || drop[id]) continue;
            drop[id] = 1;
            dropped.push_back(id);
            if (m->blk[id].isleaf) add.push_back({morton_key(m, id), id});
        }
        std::sort(add.begin(), add.end());
        std::vector<std::pair<unsigned long long, int>> &out = m->morton_next;
        out.resize(sorted.size() + add.size());
        auto *o = out.data();
        const auto *ad = add.data(), *ae = ad + add.size();
        for (const auto &e : sorted) {
            if (drop[e.second]) continue;
            while (ad != ae && *ad < e) *o++ = *ad++;
            *o++ = e;
        }
        while (ad != ae) *o++ = *ad++;
        out.resize(o - out.data());
        sorted.swap(out);
        for (int id : dropped) drop[id] = 0;
        if ((int)sorted.size() != n) m->morton_all = true;   // the leaf set disagrees: start over
    }
    if (m->morton_all) {
        sorted.resize(n);
#pragma omp parallel for schedule(static) num_threads(std::min(omp_get_max_threads(), 8)) if (n > 4096)
        for (int k = 0; k < n; k++) sorted[k] = {morton_key(m, m->leaf[k]), m->leaf[k]};
        if (n > 4096)
            __gnu_parallel::sort(sorted.begin(), sorted.end());
        else
            std::sort(sorted.begin(), sorted.end());
        m->morton_all = false;
    }
    m->morton_changed.clear();
    if (m->order == 1) {   // the Morton curve as it runs
        m->disp.resize(n);
        for (int k = 0; k < n; k++) m->disp[k] = sorted[k].second;
        return;
    }
    std::vector<std::pair<unsigned long long, int>> runs;
    const std::vector<std::pair<unsigned long long, int>> *key = &sorted;
    if (m->order == 3) {   // each run's blocks with the most ghost operand loads first (on a copy)
        runs = sorted;
        key = &runs;
    }
    const int q = n / 8, r = n % 8;   // runs 0..r-1 hold q + 1 blocks, the others q
    m->disp.assign(n, 0);
    int start = 0;
    for (int c = 0; c < 8; c++) {
        const int len = q + (c < r ? 1 : 0);
        if (m->order == 3)
            std::stable_sort(runs.begin() + start, runs.begin() + start + len,
                             [m](const std::pair<unsigned long long, int> &x, const std::pair<unsigned long long, int> &y) {
                                 return m->cost_id[x.second] > m->cost_id[y.second];
                             });
        for (int i = 0; i < len; i++) m->disp[8 * i + c] = (*key)[start + i].second;
        start += len;
    }
}

// the device table row of block n (BlockDev and its 8 NbDesc) into bd, nbt;
// a missing neighbour (2:1 balance violated) sets *bad_k
void table_row(amrlbm_mesh *m, int n, BlockDev &d, NbDesc *nbt, int *bad_k)
{
    const Block &b = m->blk[n];
    d = BlockDev{};
    for (int k = 0; k < 8; k++) nbt[k] = NbDesc{0, -1, 0, -1};
    d.level = b.level;
    d.dx = b.dx;
    d.I = b.I;
    d.J = b.J;
    d.isleaf = b.isleaf ? 1 : 0;
    d.child0 = (b.level >= 0 && !b.isleaf && b.has_child) ? b.child[0] - 1 : -1;
    d.hasleaf = (b.level >= 0 && b.hasleaf) ? 1 : 0;
    d.inst = b.level >= 0 ? inst_of(m, b) : 0;
    if (b.level < 0) return;
    for (int k = 1; k <= 8; k++) {
        const int v = nb_value(m, n, k);
        if (v == 0) {
            if (!*bad_k) *bad_k = k;
            continue;
        }
        NbDesc &e = nbt[k - 1];
        e.id = v;
        if (v > 0) {
            const Block &q = m->blk[v];
            e.level = q.level;
            e.isleaf = q.isleaf ? 1 : 0;
            e.child0 = (!q.isleaf && q.has_child) ? q.child[0] - 1 : -1;
        }
    }
}

bool same_row(const BlockDev &a, const BlockDev &b) { return std::memcmp(&a, &b, sizeof(BlockDev)) == 0; }

int upload_tables(amrlbm_mesh *m)
{
    if (!m->tables_dirty) return AMRLBM_OK;
    HostTimer timer(m->host_ms_tables);
    const int nb = m->nblks;
    // tables, lists and programs go out in one upload batch
    int rcb = batch_begin(m);
    if (rcb) return rcb;
    if ((int)m->h_bd.size() < nb) {
        m->h_bd.resize(nb);
        m->h_nbd.resize(8 * (size_t)nb);
    }
    int bad_block = 0, bad_k = 0;   // a block without neighbour k (2:1 balance violated)
    if (m->rows_all) {
        // every row, on the same team size as the program rebuild that follows (it finds the threads awake)
#pragma omp parallel for schedule(static) num_threads(nb >= 256 ? std::min(omp_get_max_threads(), 8) : 1)
        for (int n = 1; n <= nb; n++) {
            int k = 0;
            table_row(m, n, m->h_bd[n - 1], &m->h_nbd[(size_t)(n - 1) * 8], &k);
            if (k) {
#pragma omp critical(amrlbm_bad_nb)
                if (!bad_block || n < bad_block) {
                    bad_block = n;
                    bad_k = k;
                }
            }
        }
        size_t tab_at = 0;
        rcb = batch_alloc(m, sizeof(BlockDev) * nb + sizeof(NbDesc) * 8 * (size_t)nb, &tab_at);
        if (rcb) return rcb;
        std::memcpy(m->h_up + tab_at, m->h_bd.data(), sizeof(BlockDev) * nb);
        std::memcpy(m->h_up + tab_at + sizeof(BlockDev) * nb, m->h_nbd.data(), sizeof(NbDesc) * 8 * (size_t)nb);
        batch_seg(m, m->d_blk, tab_at, sizeof(BlockDev) * nb);
        batch_seg(m, m->d_nbd, tab_at + sizeof(BlockDev) * nb, sizeof(NbDesc) * 8 * (size_t)nb);
    } else {
        // the rows the regrid marked, in runs of consecutive ids
        std::sort(m->dirty_rows.begin(), m->dirty_rows.end());
        std::vector<int> &d = m->dirty_rows;
        for (int n : d) {
            if (n > nb) continue;
            int k = 0;
            table_row(m, n, m->h_bd[n - 1], &m->h_nbd[(size_t)(n - 1) * 8], &k);
            if (k && (!bad_block || n < bad_block)) {
                bad_block = n;
                bad_k = k;
            }
        }
        for (size_t i = 0; i < d.size();) {
            if (d[i] > nb) break;
            size_t j = i + 1;
            while (j < d.size() && d[j] == d[j - 1] + 1 && d[j] <= nb) j++;
            const int first = d[i] - 1, cnt = (int)(j - i);
            size_t at = 0;
            rcb = batch_alloc(m, sizeof(BlockDev) * cnt + sizeof(NbDesc) * 8 * (size_t)cnt, &at);
            if (rcb) return rcb;
            std::memcpy(m->h_up + at, &m->h_bd[first], sizeof(BlockDev) * cnt);
            std::memcpy(m->h_up + at + sizeof(BlockDev) * cnt, &m->h_nbd[(size_t)first * 8], sizeof(NbDesc) * 8 * cnt);
            batch_seg(m, m->d_blk + first, at, sizeof(BlockDev) * cnt);
            batch_seg(m, m->d_nbd + (size_t)first * 8, at + sizeof(BlockDev) * cnt, sizeof(NbDesc) * 8 * (size_t)cnt);
            i = j;
        }
        if (std::getenv("AMRLBM_VERIFY_TABLES")) {   // every row against a fresh build
            if (m->leaf_list_bad) {
                m->leaf_list_bad = false;
                return fail(AMRLBM_ERR_TOPOLOGY, "the leaf list kept across the regrid differs from a scan");
            }
            for (int n = 1; n <= nb; n++) {
                BlockDev fb;
                NbDesc fn[8];
                int k = 0;
                table_row(m, n, fb, fn, &k);
                if (!same_row(fb, m->h_bd[n - 1]) ||
                    std::memcmp(fn, &m->h_nbd[(size_t)(n - 1) * 8], sizeof(fn)) != 0)
                    return fail(AMRLBM_ERR_TOPOLOGY, "table row of block " + std::to_string(n) +
                                                         " was not marked by the regrid that changed it");
            }
        }
    }
    for (int n : m->dirty_rows) m->dirty_mark[n] = 0;
    m->leaf_delta_all = m->rows_all;
    m->leaf_delta.assign(m->dirty_rows.begin(), m->dirty_rows.end());
    if (m->rows_all) {
        m->check_all = true;
        m->changed_blocks.clear();
        m->morton_all = true;
        m->morton_changed.clear();
    } else {
        m->changed_blocks.insert(m->changed_blocks.end(), m->dirty_rows.begin(), m->dirty_rows.end());
        m->morton_changed.insert(m->morton_changed.end(), m->dirty_rows.begin(), m->dirty_rows.end());
    }
    m->dirty_rows.clear();
    m->rows_all = false;
    if (bad_block)
        return fail(AMRLBM_ERR_TOPOLOGY, "block " + std::to_string(bad_block) + " has no neighbour " +
                                             std::to_string(bad_k) + " (2:1 balance violated)");
    if (!m->leaf.empty()) {   // the programs first: launch order 3 weighs them
        int rc0 = rebuild_programs(m);
        if (rc0) return rc0;
    }
    // the leaves in launch order
    const size_t nlist = m->leaf.size();
    size_t lists_at = 0;
    rcb = batch_alloc(m, sizeof(int) * nlist, &lists_at);
    if (rcb) return rcb;
    int *lists = reinterpret_cast<int *>(m->h_up + lists_at);
    dispatch_order(m);
    for (size_t k = 0; k < nlist; k++) lists[k] = m->disp[k] - 1;
    int rc = ensure_int_buffer(&m->d_lists, &m->lists_cap, (int)nlist);
    if (rc) return rc;
    batch_seg(m, m->d_lists, lists_at, sizeof(int) * nlist);
    m->dm.blk = m->d_blk;
    m->dm.nbd = m->d_nbd;
    if (!m->leaf.empty()) {
        int rc2 = upload_programs(m);
        if (rc2) return rc2;
    }
    rc = batch_flush(m);
    if (rc) return rc;
    if (!m->leaf.empty()) {
        launch_gather_headers(m->d_phdr_slot, m->d_phdr, m->d_lists, (int)m->leaf.size(), m->stream);
        HIPCHK(hipGetLastError());
    }
    m->tables_dirty = false;
    return AMRLBM_OK;
}

const int *all_leaves(amrlbm_mesh *m) { return m->d_lists; }
int nleaves(const amrlbm_mesh *m) { return (int)m->leaf.size(); }

// The mass correction at a segment end, on the state in buffer cur: each
// instance's FP64 phase-field mass is folded and put back to its target by a
// change of C spread over the interface cells. judge / ord: the segment's
// judge and end ordinal (the launches return with the segment's end kernel).
void enqueue_mass_correction(amrlbm_mesh *m, const int *judge, int ord)
{
    const int n = nleaves(m);
    launch_mass_sum(m->dm, all_leaves(m), n, m->mass_ninst, m->d_mass_part, m->d_mass, MASS_CORRECT, 1, judge, ord,
                    m->stream);
    launch_mass_apply(m->dm, m->cur, all_leaves(m), n, m->d_mass, m->ph, judge, ord, m->stream);
}

// The budget's share of a mesh change: before restrict and prolong, the mass
// over the old leaves (MASS_SAVE); after the new tables, target += the
// difference (MASS_BUDGET). These run whatever the skip word says.
void enqueue_mass_budget(amrlbm_mesh *m, int mode)
{
    launch_mass_sum(m->dm, all_leaves(m), nleaves(m), m->mass_ninst, m->d_mass_part, m->d_mass, mode, 0, nullptr, 0,
                    m->stream);
}

int check_err_flags(amrlbm_mesh *m)
{
    int f[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(&f[0], m->d_err_flags, sizeof(int), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipMemcpyAsync(&f[1], m->d_sync + 15, sizeof(int), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    if (f[1]) {
        HIPCHK(hipMemsetAsync(m->d_sync, 0, sizeof(unsigned) * kSegSyncWords, m->stream));
        HIPCHK(hipStreamSynchronize(m->stream));
        return fail(AMRLBM_ERR_HIP, "a grid barrier of the segment kernel timed out (state undefined)");
    }
    if (f[0]) {
        HIPCHK(hipMemsetAsync(m->d_err_flags, 0, sizeof(int), m->stream));
        return fail(AMRLBM_ERR_TOPOLOGY, "ghost fill met a neighbour more than one level apart");
    }
    return AMRLBM_OK;
}

// ---------------------------------------------------------------------------
// phases: one launch over every leaf each (the ghost programs make the
// reference's coarse-to-fine level passes unnecessary)
// ---------------------------------------------------------------------------
const Prog kNoProg = {nullptr, nullptr};

int phase_fill_c(amrlbm_mesh *m, int kind, int op)
{
    int rc = upload_tables(m);
    if (rc) return rc;
    launch_prog_c(m->dm, m->cur, all_leaves(m), nleaves(m), kind, op, m->prog[kind], m->ph, m->stream);
    HIPCHK(hipGetLastError());
    return AMRLBM_OK;
}

int phase_collide(amrlbm_mesh *m)
{
    int rc = upload_tables(m);
    if (rc) return rc;
    launch_prog_c(m->dm, m->cur, all_leaves(m), nleaves(m), KIND_C2, OP_COLLIDE, kNoProg, m->ph, m->stream);
    HIPCHK(hipGetLastError());
    return AMRLBM_OK;
}

int phase_fill_hg(amrlbm_mesh *m)
{
    int rc = upload_tables(m);
    if (rc) return rc;
    launch_prog_hg(m->dm, m->cur, all_leaves(m), nleaves(m), m->prog[KIND_HG], false, false, false, m->stream);
    HIPCHK(hipGetLastError());
    return AMRLBM_OK;
}

// stream only; the ghost ring travels with the populations so that the state
// matches the reference's in-place stream cell for cell
int phase_stream(amrlbm_mesh *m, bool macro_h)
{
    int rc = upload_tables(m);
    if (rc) return rc;
    launch_prog_hg(m->dm, m->cur, all_leaves(m), nleaves(m), kNoProg, true, macro_h, true, m->stream);
    HIPCHK(hipGetLastError());
    m->cur ^= 1;
    return AMRLBM_OK;
}

int phase_macro_h(amrlbm_mesh *m)
{
    int rc = upload_tables(m);
    if (rc) return rc;
    launch_macro_h(m->dm, m->cur, all_leaves(m), nleaves(m), m->stream);
    HIPCHK(hipGetLastError());
    return AMRLBM_OK;
}

int phase_macro_g(amrlbm_mesh *m)
{
    int rc = upload_tables(m);
    if (rc) return rc;
    launch_prog_c(m->dm, m->cur, all_leaves(m), nleaves(m), KIND_C1, OP_MACRO_G, kNoProg, m->ph, m->stream);
    HIPCHK(hipGetLastError());
    return AMRLBM_OK;
}

// ---------------------------------------------------------------------------
// regrid: Refine_Derefine_Mesh (Refine_Derefine.f90:2-14)
// ---------------------------------------------------------------------------
// extra: blocks flagged for refinement besides the criterion; frozen: the
// instances of an ensemble mesh left as they are (no flags)
void clear_regrid_flags(amrlbm_mesh *m);

// Refine and Deref flags of Refine_Derefine_Mesh (Refine_Derefine.f90:18-194)
// computed on the host from Err_Interface of every leaf (downloaded here)
int host_flags(amrlbm_mesh *m, double refin_tol, double deref_tol, const std::vector<char> *extra,
               const std::vector<char> *frozen, const int *d_pred, long long *pred_count)
{
    const int nleaf = (int)m->leaf.size();
    // Err_Interface of every leaf on the device (Refine_Derefine.f90:36, :122)
    launch_err(m->dm, m->d_lists, nleaf, m->d_err, nullptr, m->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(m->h_pinned, m->d_err, sizeof(double) * nleaf, hipMemcpyDeviceToHost, m->stream));
    std::vector<int> pred(d_pred ? m->nblks : 0);
    if (d_pred) HIPCHK(hipMemcpyAsync(pred.data(), d_pred, sizeof(int) * m->nblks, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    std::vector<double> err(m->nblks + 1, 0.0);
    for (int k = 0; k < nleaf; k++) err[m->disp[k]] = m->h_pinned[k];

    auto is_frozen = [&](const Block &b) { return frozen && (*frozen)[inst_of(m, b)]; };
    // Flag_for_refinement (Refine_Derefine.f90:18-54)
    for (int n = 1; n <= m->nblks; n++) {
        Block &b = m->blk[n];
        b.refin = false;
        if (b.isleaf && !is_frozen(b)) {
            const int lo = min_level_of(m, b), hi = max_level_of(m, b);
            if (b.level < lo) b.refin = true;
            else if (b.level < hi && err[n] > refin_tol) b.refin = true;
            else if (b.level < hi && ((extra && (*extra)[n]) || (d_pred && pred[n - 1]))) {   // predicted
                b.refin = true;
                if (pred_count) ++*pred_count;
            }
        }
    }
    // Enforce_1_Level_Jump / Check_Level_Jump (:58-93): least fixed point
    std::vector<int> work;
    for (int n = 1; n <= m->nblks; n++)
        if (m->blk[n].refin) work.push_back(n);
    while (!work.empty()) {
        int n = work.back();
        work.pop_back();
        for (int k = 1; k <= 8; k++) {
            int b = nb_value(m, n, k);
            if (b > 0 && !m->blk[b].refin && m->blk[b].level == m->blk[n].level - 1 && m->blk[b].isleaf) {
                m->blk[b].refin = true;
                work.push_back(b);
            }
        }
    }
    // Flag_for_derefinement (:97-145)
    for (int n = 1; n <= m->nblks; n++) {
        Block &b = m->blk[n];
        b.deref = false;
        if (!b.hasleaf || b.level < min_level_of(m, b) || is_frozen(b)) continue;
        int c1 = b.child[0];
        bool any = false;
        for (int c = c1; c <= c1 + 3; c++) any = any || m->blk[c].refin;
        if (any) continue;
        b.deref = true;
        for (int c = c1; c <= c1 + 3; c++)
            if (err[c] > deref_tol) {
                b.deref = false;
                break;
            }
    }
    // enforce_1_Level_jump_deref_2D / Check_Level_Jump_Deref_2D (:149-194)
    static const int pair_c1[8] = {2, 1, 3, 1, 4, 2, 3, 1};
    static const int pair_c2[8] = {4, 3, 4, 2, 4, 2, 3, 1};
    for (int L = m->levelmax - 1; L >= m->levelmin - 1; L--)
        for (int n = 1; n <= m->nblks; n++) {
            if (m->blk[n].level != L || !m->blk[n].deref) continue;
            for (int k = 1; k <= 8 && m->blk[n].deref; k++) {
                int b = nb_value(m, n, k);
                if (b > 0 && !m->blk[b].isleaf) {
                    int cc[2] = {m->blk[b].child[pair_c1[k - 1] - 1], m->blk[b].child[pair_c2[k - 1] - 1]};
                    for (int c : cc)
                        if (m->blk[c].refin || (!m->blk[c].isleaf && !m->blk[c].deref)) {
                            m->blk[n].deref = false;
                            break;
                        }
                }
            }
        }
    return AMRLBM_OK;
}

// d_pred: the learned predictor's flags by block slot on the device (read with
// Err_Interface); *pred_count grows by the leaves refined on them alone;
// dev_flags: the Refine and Deref flags already decided on the device (the
// flag lists of k_regrid_check), used instead of the host's
int regrid(amrlbm_mesh *m, double refin_tol, double deref_tol, const std::vector<char> *extra = nullptr,
           const std::vector<char> *frozen = nullptr, const int *d_pred = nullptr, long long *pred_count = nullptr,
           const int *dev_flags = nullptr)
{
    int rc = upload_tables(m);
    if (rc) return rc;
    HostTimer timer(m->host_ms_regrid);
    if (dev_flags) {
        // the device decision's flags (k_regrid_check), as host_flags sets them
        clear_regrid_flags(m);
        for (int i = 0; i < dev_flags[1]; i++) m->blk[dev_flags[3 + i] + 1].refin = true;
        for (int i = 0; i < dev_flags[2]; i++) m->blk[dev_flags[kRegridDerefAt + i] + 1].deref = true;
        for (int i = 0; i < dev_flags[1]; i++) m->flagged.push_back(dev_flags[3 + i] + 1);
        for (int i = 0; i < dev_flags[2]; i++) m->flagged.push_back(dev_flags[kRegridDerefAt + i] + 1);
        if (std::getenv("AMRLBM_VERIFY_FLAGS")) {
            std::vector<char> dev(2 * (m->nblks + 1));
            for (int n = 1; n <= m->nblks; n++) {
                dev[2 * n] = m->blk[n].refin;
                dev[2 * n + 1] = m->blk[n].deref;
            }
            if ((rc = host_flags(m, refin_tol, deref_tol, nullptr, nullptr, nullptr, nullptr))) return rc;
            for (int n = 1; n <= m->nblks; n++)
                if (dev[2 * n] != m->blk[n].refin || dev[2 * n + 1] != m->blk[n].deref)
                    return fail(AMRLBM_ERR_STATE, "device regrid flags of block " + std::to_string(n) +
                                                      " differ from the host's");
        }
    } else {
        if ((rc = host_flags(m, refin_tol, deref_tol, extra, frozen, d_pred, pred_count))) return rc;
        m->flagged.clear();
        for (int n = 1; n <= m->nblks; n++)
            if (m->blk[n].refin || m->blk[n].deref) m->flagged.push_back(n);
    }
    // the flagged blocks in the reference's loop orders: Derefine by ascending
    // id (the freed children are unflagged leaves, so freeing never flags or
    // unflags a later block), Refine level by level, each level by ascending
    // id (the children it creates are unflagged, so the passes over every id
    // would find the same blocks in the same order)
    std::vector<int> dlist, rlist;
    for (int n : m->flagged) {
        if (m->blk[n].deref) dlist.push_back(n);
        if (m->blk[n].refin && m->blk[n].level >= m->levelmin && m->blk[n].level <= m->levelmax) rlist.push_back(n);
    }
    std::sort(dlist.begin(), dlist.end());
    std::sort(rlist.begin(), rlist.end(), [m](int a, int b) {
        return m->blk[a].level != m->blk[b].level ? m->blk[a].level < m->blk[b].level : a < b;
    });
    // Derefine (:198-211): restrict on the device, then free the children
    std::vector<int> rpairs;
    for (int n : dlist) {
        rpairs.push_back(n - 1);
        rpairs.push_back(m->blk[n].child[0] - 1);
        m->blk[n].alloc = true;
        remove_children(m, n);
    }
    // Refine (:262-294): ids from the free-list head or Nblks+1
    std::vector<int> ppairs;
    {
        for (int n : rlist) {
            if (!m->blk[n].refin) continue;
            int c1;
            if (!m->freeq.empty()) {
                c1 = m->freeq.front();
                rc = add_child_blocks(m, n, c1);
                if (rc) return rc;
                m->kk -= 4;
                m->freeq.pop_front();
            } else {
                c1 = m->nblks + 1;
                rc = add_child_blocks(m, n, c1);
                if (rc) return rc;
            }
            m->blk[n].alloc = false;
            ppairs.push_back(n - 1);
            ppairs.push_back(c1 - 1);
        }
    }
    // device data movement: restrict first (freed slots may be reused by refine)
    std::vector<int> pairs(rpairs);
    pairs.insert(pairs.end(), ppairs.begin(), ppairs.end());
    if (!pairs.empty()) {
        rc = ensure_int_buffer(&m->d_pairs, &m->pairs_cap, (int)pairs.size());
        if (rc) return rc;
        // pinned, rewritten once the previous copy out of it has run
        if (m->pairs_recorded) HIPCHK(hipEventSynchronize(m->ev_pairs));
        if (pairs.size() > m->hpairs_cap) {
            if (m->h_pairs) HIPCHK(hipHostFree(m->h_pairs));
            m->h_pairs = nullptr;
            m->hpairs_cap = pairs.size() + pairs.size() / 2 + 1024;
            HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&m->h_pairs), sizeof(int) * m->hpairs_cap, hipHostMallocDefault));
        }
        std::memcpy(m->h_pairs, pairs.data(), sizeof(int) * pairs.size());
        HIPCHK(hipMemcpyAsync(m->d_pairs, m->h_pairs, sizeof(int) * pairs.size(), hipMemcpyHostToDevice, m->stream));
        HIPCHK(hipEventRecord(m->ev_pairs, m->stream));
        m->pairs_recorded = true;
        // the kernels read the pre-regrid block table (parent dx, children);
        // it is replaced only after them (the old leaf list too, which the
        // mass budget sums over first)
        if (m->mass_on) enqueue_mass_budget(m, MASS_SAVE);
        launch_restrict(m->dm, m->cur, m->d_pairs, (int)rpairs.size() / 2, m->stream);
        launch_prolong(m->dm, m->cur, m->d_pairs + rpairs.size(), (int)ppairs.size() / 2, m->ph, m->stream);
        HIPCHK(hipGetLastError());
        // no wait here: the table upload below is ordered after these kernels on the stream
    }
    if (rpairs.empty() && ppairs.empty()) return AMRLBM_OK;   // flags only: tables and graphs stay valid
    for (size_t i = 0; i < rpairs.size(); i += 2) {
        mark_changed(m, rpairs[i] + 1);
        for (int c = rpairs[i + 1] + 1; c <= rpairs[i + 1] + 4; c++) mark_row(m, c);   // freed (or reused below)
    }
    for (size_t i = 0; i < ppairs.size(); i += 2) {
        mark_changed(m, ppairs[i] + 1);
        for (int c = ppairs[i + 1] + 1; c <= ppairs[i + 1] + 4; c++) mark_changed(m, c);
    }
    array_of_leaf_blocks(m, true);
    timer.stop();
    rc = upload_tables(m);
    if (rc || !m->mass_on) return rc;
    enqueue_mass_budget(m, MASS_BUDGET);
    HIPCHK(hipGetLastError());
    return AMRLBM_OK;
}

// ---------------------------------------------------------------------------
// layout conversion between the reference's h(9,0:nx+1,0:ny+1) and the
// device's [set][q][y][x]
// ---------------------------------------------------------------------------
// element size of the mesh's fields and the address of element `off` of one
size_t esz(const amrlbm_mesh *m) { return m->dm.f32 ? sizeof(float) : sizeof(double); }
void *fptr(const amrlbm_mesh *m, void *base, size_t off) { return static_cast<char *>(base) + off * esz(m); }

// FP32 storage shifts: h_q - w_q, C - 1 (deviations from the liquid at rest)
const double kW[9] = {16.0 / 36.0, 4.0 / 36.0, 4.0 / 36.0, 4.0 / 36.0, 4.0 / 36.0,
                      1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};

// n values at element offset off of a device field to host doubles, adding
// shift(k) (k = element index) on an FP32 mesh
template <class Shift>
int download(amrlbm_mesh *m, void *base, size_t off, size_t n, double *out, Shift shift)
{
    if (!m->dm.f32) {
        HIPCHK(hipMemcpyAsync(out, fptr(m, base, off), sizeof(double) * n, hipMemcpyDeviceToHost, m->stream));
        HIPCHK(hipStreamSynchronize(m->stream));
        return AMRLBM_OK;
    }
    std::vector<float> f(n);
    HIPCHK(hipMemcpyAsync(f.data(), fptr(m, base, off), sizeof(float) * n, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    for (size_t k = 0; k < n; k++) out[k] = (double)f[k] + shift(k);
    return AMRLBM_OK;
}

template <class Shift>
int upload(amrlbm_mesh *m, void *base, size_t off, size_t n, const double *in, Shift shift)
{
    if (!m->dm.f32) {
        HIPCHK(hipMemcpyAsync(fptr(m, base, off), in, sizeof(double) * n, hipMemcpyHostToDevice, m->stream));
        HIPCHK(hipStreamSynchronize(m->stream));
        return AMRLBM_OK;
    }
    std::vector<float> f(n);
    for (size_t k = 0; k < n; k++) f[k] = (float)(in[k] - shift(k));
    HIPCHK(hipMemcpyAsync(fptr(m, base, off), f.data(), sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return AMRLBM_OK;
}

auto no_shift = [](size_t) { return 0.0; };

void ref_to_dev_pop(const double *ref, double *dev)   // one set
{
    for (int y = 0; y < NY + 2; y++)
        for (int x = 0; x < NX + 2; x++)
            for (int q = 0; q < 9; q++) dev[q * GS + pcell(x, y)] = ref[q + 9 * (x + FW * y)];
}

void dev_to_ref_pop(const double *dev, double *ref)
{
    for (int y = 0; y < NY + 2; y++)
        for (int x = 0; x < NX + 2; x++)
            for (int q = 0; q < 9; q++) ref[q + 9 * (x + FW * y)] = dev[q * GS + pcell(x, y)];
}

int alloc_device(amrlbm_mesh *m)
{
    const size_t S = (size_t)m->prm.max_blocks, e = esz(m);
    HIPCHK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    HIPCHK(hipMalloc(&m->dm.pop[0], e * POPB * S));
    HIPCHK(hipMalloc(&m->dm.pop[1], e * POPB * S));
    HIPCHK(hipMalloc(&m->dm.C, e * CS * S));
    HIPCHK(hipMalloc(&m->dm.PU, e * PUB * S));
    HIPCHK(hipMemsetAsync(m->dm.pop[0], 0, e * POPB * S, m->stream));
    HIPCHK(hipMemsetAsync(m->dm.pop[1], 0, e * POPB * S, m->stream));
    HIPCHK(hipMemsetAsync(m->dm.C, 0, e * CS * S, m->stream));
    HIPCHK(hipMemsetAsync(m->dm.PU, 0, e * PUB * S, m->stream));
    HIPCHK(hipMalloc(&m->d_blk, sizeof(BlockDev) * S));
    HIPCHK(hipMalloc(&m->d_nbd, sizeof(NbDesc) * 8 * S));
    for (int kind = 0; kind < 3; kind++) HIPCHK(hipMalloc(&m->d_phdr_slot[kind], sizeof(PHdr) * S));
    HIPCHK(hipMalloc(&m->d_sync, sizeof(unsigned) * kSegSyncWords));
    HIPCHK(hipMalloc(&m->d_judge, sizeof(int)));
    HIPCHK(hipMemsetAsync(m->d_sync, 0, sizeof(unsigned) * kSegSyncWords, m->stream));
    HIPCHK(hipMalloc(&m->d_err, sizeof(double) * S));
    HIPCHK(hipMalloc(&m->d_err_slot, sizeof(double) * S));
    HIPCHK(hipMalloc(&m->d_changed, sizeof(int) * (4 + kRegridFlagWords)));
    HIPCHK(hipMemsetAsync(m->d_changed, 0, sizeof(int) * (4 + kRegridFlagWords), m->stream));
    m->d_err_flags = m->d_changed + 3;
    HIPCHK(hipHostMalloc(&m->h_changed, sizeof(int) * (4 + kRegridFlagWords), hipHostMallocDefault));
    HIPCHK(hipEventCreate(&m->ev_check));   // timed: a stepping call that ends on a decision ends its timing there
    HIPCHK(hipEventCreateWithFlags(&m->ev_tab, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&m->ev_pairs, hipEventDisableTiming));
    HIPCHK(hipMalloc(&m->d_part, sizeof(double) * 9 * S));
    HIPCHK(hipMalloc(&m->d_diag, sizeof(double) * 16));
    HIPCHK(hipHostMalloc(&m->h_pinned, sizeof(double) * std::max<size_t>(S, 16), hipHostMallocDefault));
    HIPCHK(hipEventCreate(&m->ev0));
    HIPCHK(hipEventCreate(&m->ev1));
    HIPCHK(hipStreamSynchronize(m->stream));
    m->dm.err_flags = m->d_err_flags;
    m->dm.skip = m->d_changed;
    m->dm.blk = m->d_blk;
    m->dm.nbd = m->d_nbd;
    return AMRLBM_OK;
}

void reset_topology(amrlbm_mesh *m)
{
    m->blk.assign((size_t)m->prm.max_blocks + 1, Block());
    m->at.assign(m->nlev(), std::vector<int>());
    for (int L = 0; L < m->nlev(); L++) m->at[L].assign((size_t)m->width(L) * m->height(L), 0);
    m->nblks = 0;
    m->kk = 0;
    m->T = 0;
    m->freeq.clear();
    m->leaf.clear();
    m->tables_dirty = true;
    m->rows_all = true;
    m->topo_version++;
    m->mass_on = 0;   // a new state has no budget: the caller switches the correction on again
}

// analytic phase field of Initialize_AMR_Rising_Bubble (AMR_Rising_2D_OMP.f90:174-185)
void bubble_c(const amrlbm_mesh *m, const Block &b, double R, double X0, double Y0, double W, double *C)
{
    const int dx = b.dx;
    const int J = b.J % m->tile_height(b.level);   // the instance's own frame
    for (int j = -1; j <= NY + 2; j++)
        for (int i = -1; i <= NX + 2; i++) {
            double Xc = (double)(b.I * dx * NX) + (double)dx * ((double)i - 0.5);
            double Yc = (double)(J * dx * NY) + (double)dx * ((double)j - 0.5);
            double Ri = std::sqrt((Xc - X0) * (Xc - X0) + (Yc - Y0) * (Yc - Y0));
            C[(j + 1) * CW + (i + 1)] = 0.5 + 0.5 * std::tanh(2.0 * (Ri - R) / W);
        }
}

int init_bubble(amrlbm_mesh *m, double R, double X0, double Y0)
{
    int rc = upload_tables(m);
    if (rc) return rc;
    const size_t nl = m->leaf.size();
    std::vector<double> cbuf(nl * CS);
    for (size_t k = 0; k < nl; k++) bubble_c(m, m->blk[m->leaf[k]], R, X0, Y0, m->ph.W, &cbuf[k * CS]);
    const double cs = m->dm.f32 ? 1.0 : 0.0;
    for (size_t k = 0; k < nl; k++) {
        int rc2 = upload(m, m->dm.C, (size_t)(m->leaf[k] - 1) * CS, CS, &cbuf[k * CS], [cs](size_t) { return cs; });
        if (rc2) return rc2;
    }
    launch_init_fields(m->dm, m->cur, m->d_lists, (int)nl, m->ph, R, m->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(m->stream));
    return AMRLBM_OK;
}

// A segment of len whole steps, starting from a state whose U and P are
// current (after set-up, a regrid or a previous segment) and ending with U
// and P current again, two launches per step over every leaf:
//   step 1      : [C2 fill + collide]          [hg fill + stream + C = sum h]
//   steps 2..len: [C2 fill + U,P + collide]    [hg fill + stream + C = sum h]
//   end         : [C fill + U,P]
// The collision writes the other population buffer and the stream writes
// back, so the state stays in buffer m->cur from step to step, and a step
// that does not run leaves it whole. With judge, every state after the
// first is tested against the plain refinement criterion in its collision
// kernel (the end kernel tests the last): the first state found needing
// refinement stops the segment there (d_judge = its ordinal, the later
// launches return at once), so the regrid can follow at once.
// after(kind, cells) is called after every launch (kind 0: C2 kernel,
// 1: hg kernel, 2: C kernel).
template <class After>
int enqueue_segment(amrlbm_mesh *m, int len, bool judge, After after)
{
    const int n = nleaves(m);
    const int *lst = all_leaves(m);
    const int cur = m->cur;
    StepCtl ctl;
    ctl.out = cur ^ 1;
    if (judge) {
        ctl.judge = m->d_judge;
        ctl.min_level = m->prm.min_levels;
        ctl.max_level = m->prm.max_levels;
        ctl.refin_tol = 0.002;
        launch_fill_int(m->d_judge, INT_MAX, m->stream);
    }
    for (int s = 0; s < len; s++) {
        ctl.ord = s;
        ctl.judge_on = judge && s > 0;   // the state at a segment's start was decided on before it
        launch_prog_c(m->dm, cur, lst, n, KIND_C2, s == 0 ? OP_COLLIDE : OP_MACRO_G_COLLIDE, m->prog[KIND_C2], m->ph,
                      m->stream, nullptr, ctl);
        int rc = after(0, n * PS);
        if (rc) return rc;
        launch_prog_hg(m->dm, cur ^ 1, lst, n, m->prog[KIND_HG], true, true, false, m->stream, ctl.judge, s + 1);
        rc = after(1, n * PS);
        if (rc) return rc;
    }
    // Err_Interface of every leaf comes out of the last kernel, for a regrid
    // decision that may follow the segment
    StepCtl end = ctl;
    end.out = -1;
    end.ord = len;
    end.judge_on = judge;
    // the mass correction before the end kernel, so that U, P and
    // Err_Interface follow the corrected C
    if (m->mass_on) enqueue_mass_correction(m, ctl.judge, len);
    launch_prog_c(m->dm, cur, lst, n, KIND_C1, OP_MACRO_G, m->prog[KIND_C1], m->ph, m->stream, m->d_err_slot, end);
    return after(2, n * PS);
}

void drop_graphs(amrlbm_mesh *m)
{
    for (auto &g : m->graphs) (void)hipGraphExecDestroy(g.exec);
    m->graphs.clear();
}

// The executable graph of a segment of len steps from the current buffer
// parity. After a topology change the segment has the same launches with new
// grids and arguments, so the new capture updates the existing executable
// (hipGraphExecUpdate) instead of instantiating another; an update the
// runtime refuses falls back to instantiation.
int segment_graph(amrlbm_mesh *m, int len, hipGraphExec_t *out, bool judge = false)
{
    amrlbm_mesh::GraphEntry *entry = nullptr;
    for (auto &g : m->graphs)
        if (g.len == len && g.parity == m->cur && g.judge == judge) entry = &g;
    if (entry && entry->topo == m->topo_version) {
        *out = entry->exec;
        return AMRLBM_OK;
    }
    HostTimer timer(m->host_ms_graphs);
    m->graphs_built++;
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_segment(m, len, judge, [](int, int) { return AMRLBM_OK; });
    hipError_t e = hipStreamEndCapture(m->stream, &g);
    if (rc) return rc;
    if (e != hipSuccess) return fail(AMRLBM_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    if (entry) {
        hipGraphNode_t bad = nullptr;
        hipGraphExecUpdateResult res = hipGraphExecUpdateError;
        if (hipGraphExecUpdate(entry->exec, g, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
            HIPCHK(hipGraphDestroy(g));
            entry->topo = m->topo_version;
            *out = entry->exec;
            return AMRLBM_OK;
        }
        (void)hipGetLastError();   // a refused update leaves an error code behind
        HIPCHK(hipGraphExecDestroy(entry->exec));
    }
    hipGraphExec_t ex;
    HIPCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
    if (entry) {
        entry->exec = ex;
        entry->topo = m->topo_version;
    } else {
        m->graphs.push_back({len, m->cur, judge, ex, m->topo_version});
    }
    *out = ex;
    return AMRLBM_OK;
}

// Whether segments run as the persistent segment kernel: chosen (exec 1) and
// every leaf's workgroup fits on the GPU at once with the LDS the current
// ghost programs need.
bool use_segment_kernel(amrlbm_mesh *m)
{
    if (m->exec != 1 || m->mass_on) return false;   // the segment kernel carries no mass correction
    const size_t lds = segment_lds_bytes(m->dm, m->prog[KIND_HG]);
    if (lds != m->seg_lds) {
        m->seg_lds = lds;
        m->seg_cap = segment_capacity(m->dm, m->prog[KIND_HG]);
    }
    return nleaves(m) <= m->seg_cap;
}

// One segment of len steps as one launch of the persistent segment kernel
// (the same state afterwards as the segment's step graph).
int launch_segment_kernel(amrlbm_mesh *m, int len)
{
    m->seg_kernel_ran = true;
    hipError_t e = launch_segment(m->dm, m->cur, all_leaves(m), nleaves(m), m->prog[KIND_C2], m->prog[KIND_HG],
                                  m->prog[KIND_C1], m->ph, len, m->d_err_slot, m->d_sync, m->stream);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(AMRLBM_ERR_HIP, std::string("segment kernel launch: ") + hipGetErrorString(e));
    }
    return AMRLBM_OK;
}

// The segment of len steps from the current state, as the tuning says: the
// persistent segment kernel, or the segment's HIP graph (always the graph
// with the judge, which the segment kernel does not carry).
int launch_steps(amrlbm_mesh *m, int len, bool judge = false)
{
    if (!judge && use_segment_kernel(m)) return launch_segment_kernel(m, len);
    hipGraphExec_t ex;
    int rc = segment_graph(m, len, &ex, judge);
    if (rc) return rc;
    HIPCHK(hipGraphLaunch(ex, m->stream));
    return AMRLBM_OK;
}

// Enqueue the device regrid decision at the current step, behind a segment
// whose last kernel left Err_Interface of every leaf in d_err_slot:
// k_regrid_check writes 1 to d_changed iff the host regrid would change the
// mesh (or it cannot tell) and the flags it would set after it, and their
// copy reaches h_changed when ev_check fires.
// Until the next check, a nonzero d_changed makes the step kernels return at
// once, so a segment launched behind the check leaves the state as it is.
int enqueue_regrid_check(amrlbm_mesh *m)
{
    HIPCHK(launch_regrid_check(m->dm, m->nblks, m->d_err_slot, m->prm.min_levels, m->prm.max_levels, m->levelmin,
                               m->levelmax, 0.002, 0.001, m->d_changed, m->stream, nullptr, 0, m->d_changed + 4));
    // the decision and its flag lists in one copy
    HIPCHK(hipMemcpyAsync(m->h_changed, m->d_changed, sizeof(int) * (4 + kRegridFlagWords), hipMemcpyDeviceToHost,
                          m->stream));
    HIPCHK(hipEventRecord(m->ev_check, m->stream));
    m->regrid_checks++;
    return AMRLBM_OK;
}

// the flags a regrid that changes nothing leaves: Refin and Deref all false
// (those the last flag pass set)
void clear_regrid_flags(amrlbm_mesh *m)
{
    for (int n : m->flagged) m->blk[n].refin = m->blk[n].deref = false;
    m->flagged.clear();
}

// steps to the next regrid (or to the end of the call)
int segment_length(const amrlbm_mesh *m, int remaining, int regrid_every)
{
    int to_regrid = regrid_every - (m->T % regrid_every);
    return std::min(remaining, to_regrid);
}

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

const char *amrlbm_last_error(void) { return g_last_error.c_str(); }

// the kernels' constants from the module variables (D2Q9_Mod.f90:10-16); the
// quotients with the IEEE divisions the kernels would use
static Phys phys_from(const amrlbm_params *p)
{
    Phys ph{};
    ph.M0 = p->M0;
    ph.wc = 1.0 / (0.5 + p->M0 * 3.0);
    ph.W = p->W;
    ph.sigma = p->sigma;
    ph.beta = p->beta;
    ph.kappa = p->kappa;
    ph.taul = p->taul;
    ph.tauh = p->tauh;
    ph.rhol = p->rhol;
    ph.rhoh = p->rhoh;
    ph.gy = p->gy;
    ph.drho3 = (p->rhoh - p->rhol) / 3.0;
    ph.drho6 = (p->rhoh - p->rhol) / 6.0;
    ph.inv_tauh = 1.0 / p->tauh;
    ph.inv_taul = 1.0 / p->taul;
    ph.inv_W = 1.0 / p->W;
    return ph;
}

static PhysF physf_from(const amrlbm_params *p)
{
    PhysF f{};
    f.M0 = (float)p->M0;
    f.wc = (float)(1.0 / (0.5 + p->M0 * 3.0));
    f.beta = (float)p->beta;
    f.kappa = (float)p->kappa;
    f.taul = (float)p->taul;
    f.tauh = (float)p->tauh;
    f.rhol = (float)p->rhol;
    f.rhoh = (float)p->rhoh;
    f.gy = (float)p->gy;
    f.drho = (float)(p->rhoh - p->rhol);
    f.drho3 = (float)((p->rhoh - p->rhol) / 3.0);
    f.drho6 = (float)((p->rhoh - p->rhol) / 6.0);
    f.inv_tauh = (float)(1.0 / p->tauh);
    f.inv_taul = (float)(1.0 / p->taul);
    f.k12W = (float)(12.0 * p->M0 / p->W);
    return f;
}

static int create_mesh(const amrlbm_params *p, bool f32, amrlbm_mesh **out)
{
    if (!p || !out) return fail(AMRLBM_ERR_ARG, "null argument");
    *out = nullptr;
    if (p->nx != NX || p->ny != NY) return fail(AMRLBM_ERR_ARG, "this build supports 16 x 16 blocks only");
    if (p->nroots < 1 || p->min_levels < 0 || p->max_levels < p->min_levels || p->max_levels > 14 ||
        p->max_blocks < p->nroots || p->L0 != (NX << p->max_levels))
        return fail(AMRLBM_ERR_ARG, "inconsistent mesh parameters (L0 must be nx * 2^max_levels)");
    amrlbm_mesh *m = new amrlbm_mesh();
    m->prm = *p;
    m->tile_roots = p->nroots;
    m->ph = phys_from(p);
    m->dm.f32 = f32 ? 1 : 0;
    m->dm.phf = physf_from(p);
    int rc = alloc_device(m);
    if (!rc) rc = scratch_on_current_device(&m->scratch);
    if (rc) {
        amrlbm_destroy(m);
        return rc;
    }
    reset_topology(m);
    *out = m;
    return AMRLBM_OK;
}

int amrlbm_create(const amrlbm_params *p, amrlbm_mesh **out) { return create_mesh(p, false, out); }

int amrlbm_create_fp32(const amrlbm_params *p, amrlbm_mesh **out) { return create_mesh(p, true, out); }

int amrlbm_destroy(amrlbm_mesh *m)
{
    if (!m) return AMRLBM_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    drop_graphs(m);
    (void)hipFree(m->dm.pop[0]);
    (void)hipFree(m->dm.pop[1]);
    (void)hipFree(m->dm.C);
    (void)hipFree(m->dm.PU);
    (void)hipFree(m->d_blk);
    (void)hipFree(m->d_nbd);
    for (int k = 0; k < 3; k++) {
        (void)hipFree(m->d_phdr[k]);
        (void)hipFree(m->d_phdr_slot[k]);
        (void)hipFree(m->d_pnode[k]);
    }
    if (m->h_up) (void)hipHostFree(m->h_up);
    (void)hipFree(m->d_up);
    if (m->h_pairs) (void)hipHostFree(m->h_pairs);
    (void)hipFree(m->d_sync);
    (void)hipFree(m->d_judge);
    (void)hipFree(m->d_net_w);
    (void)hipFree(m->d_net_x);
    (void)hipFree(m->d_net_logit);
    (void)hipFree(m->d_pred);
    (void)hipFree(m->d_lists);
    (void)hipFree(m->d_pairs);
    (void)hipFree(m->d_err);
    (void)hipFree(m->d_err_slot);
    (void)hipFree(m->d_changed);
    if (m->h_changed) (void)hipHostFree(m->h_changed);
    if (m->ev_check) (void)hipEventDestroy(m->ev_check);
    if (m->ev_tab) (void)hipEventDestroy(m->ev_tab);
    if (m->ev_pairs) (void)hipEventDestroy(m->ev_pairs);
    (void)hipFree(m->d_part);
    (void)hipFree(m->d_diag);
    (void)hipFree(m->d_phs);
    (void)hipFree(m->d_phsf);
    (void)hipFree(m->d_lv);
    (void)hipFree(m->d_iR);
    (void)hipFree(m->d_diag_inst);
    (void)hipFree(m->d_mass);
    (void)hipFree(m->d_mass_part);
    if (m->h_pinned) (void)hipHostFree(m->h_pinned);
    if (m->ev0) (void)hipEventDestroy(m->ev0);
    if (m->ev1) (void)hipEventDestroy(m->ev1);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return AMRLBM_OK;
}

int amrlbm_setup_rising_bubble(amrlbm_mesh *m, double R, double x0, double y0)
{
    return amrlbm_setup_replicas(m, 1, R, x0, y0);
}

}  // extern "C"

// AMR_LBM_Rising_Bubble set-up (AMR_Rising_2D_OMP.f90:51-67) with init(m)
// setting the fields of every leaf (Initialize_AMR_Rising_Bubble)
static void clear_ensemble(amrlbm_mesh *m);

template <class Init>
static int setup_mesh(amrlbm_mesh *m, int ninst, Init init)
{
    if (!m) return fail(AMRLBM_ERR_ARG, "null mesh");
    if (ninst < 1 || m->prm.nroots % ninst)
        return fail(AMRLBM_ERR_ARG, "the roots must split evenly into the instances");
    clear_ensemble(m);
    m->tile_roots = m->prm.nroots / ninst;
    reset_topology(m);
    m->cur = 0;
    const amrlbm_params &p = m->prm;
    m->levelmin = m->levelmax = p.min_levels;
    // Initialize_Base_Block (AMR_Rising_2D_OMP.f90:79-115), roots stacked in y
    for (int N = 1; N <= p.nroots; N++) {
        Block &b = m->blk[N];
        b.level = 0;
        b.dx = p.L0 / p.nx;
        b.I = 0;
        b.J = N - 1;
        b.has_child = true;
        for (int k = 0; k < 4; k++) b.child[k] = -1;
        b.parent = -1;
        b.isleaf = true;
        m->nblks = N;
        place(m, N);
    }
    // Initial_Mesh (AMR_Rising_2D_OMP.f90:493-513); device slots start zeroed
    HIPCHK(hipMemsetAsync(m->dm.pop[0], 0, esz(m) * POPB * p.max_blocks, m->stream));
    HIPCHK(hipMemsetAsync(m->dm.pop[1], 0, esz(m) * POPB * p.max_blocks, m->stream));
    HIPCHK(hipMemsetAsync(m->dm.C, 0, esz(m) * CS * p.max_blocks, m->stream));
    HIPCHK(hipMemsetAsync(m->dm.PU, 0, esz(m) * PUB * p.max_blocks, m->stream));
    for (int L = 1; L <= m->levelmin; L++) {
        int nend = m->nblks;
        for (int n = 1; n <= nend; n++)
            if (m->blk[n].level == L - 1) {
                int rc = add_child_blocks(m, n, m->nblks + 1);
                if (rc) return rc;
            }
    }
    array_of_leaf_blocks(m);
    int rc = init(m);
    if (rc) return rc;
    while (m->levelmax < p.max_levels) {   // AMR_Rising_2D_OMP.f90:63-67
        rc = regrid(m, 0.002, 0.001);
        if (rc) return rc;
        rc = init(m);
        if (rc) return rc;
        m->levelmax++;
    }
    m->T = 0;
    determine_min_max_levels(m);
    m->ready = true;
    return check_err_flags(m);
}

extern "C" {

int amrlbm_setup_replicas(amrlbm_mesh *m, int ninst, double R, double x0, double y0)
{
    return setup_mesh(m, ninst, [=](amrlbm_mesh *mm) { return init_bubble(mm, R, x0, y0); });
}

int amrlbm_setup_bubbles(amrlbm_mesh *m, int n, const double *R, const double *x0, const double *y0)
{
    if (!m || n < 1 || !R || !x0 || !y0) return fail(AMRLBM_ERR_ARG, "bad bubble arguments");
    std::vector<double> bub(3 * (size_t)n);
    for (int k = 0; k < n; k++) {
        if (!(R[k] > 0.0)) return fail(AMRLBM_ERR_ARG, "bubble radius must be positive");
        bub[3 * k] = x0[k];
        bub[3 * k + 1] = y0[k];
        bub[3 * k + 2] = R[k];
    }
    DeviceBuffer buf;
    HIPCHK(hipMalloc(&buf.p, sizeof(double) * bub.size()));
    const double *d_bub = static_cast<const double *>(buf.p);
    HIPCHK(hipMemcpy(buf.p, bub.data(), sizeof(double) * bub.size(), hipMemcpyHostToDevice));
    return setup_mesh(m, 1, [&](amrlbm_mesh *mm) {
        int r2 = upload_tables(mm);
        if (r2) return r2;
        launch_init_fields(mm->dm, mm->cur, mm->d_lists, nleaves(mm), mm->ph, 0.0, mm->stream, d_bub, n);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(mm->stream));
        return AMRLBM_OK;
    });
}

}  // extern "C"

// An ensemble mesh goes back to one fluid and level range (a set-up or
// topology of another kind replaces it).
static void clear_ensemble(amrlbm_mesh *m)
{
    m->ens = 0;
    m->inst.clear();
    m->dm.phs = nullptr;
    m->dm.phsf = nullptr;
    m->dm.lv = nullptr;
    m->dm.tile_roots = 0;
}

// Fields of every leaf of an ensemble mesh (Initialize_AMR_Rising_Bubble,
// AMR_Rising_2D_OMP.f90:142-231): C from each instance's bubble and W on the
// host, as init_bubble, then U, P, h, g with the instance's constants.
static int init_ensemble(amrlbm_mesh *m)
{
    int rc = upload_tables(m);
    if (rc) return rc;
    const size_t nl = m->leaf.size();
    std::vector<double> cbuf(nl * CS);
    for (size_t k = 0; k < nl; k++) {
        const Block &b = m->blk[m->leaf[k]];
        const amrlbm_instance &q = m->inst[inst_of(m, b)];
        bubble_c(m, b, q.R, q.x0, q.y0, q.W, &cbuf[k * CS]);
    }
    const double cs = m->dm.f32 ? 1.0 : 0.0;
    for (size_t k = 0; k < nl; k++) {
        rc = upload(m, m->dm.C, (size_t)(m->leaf[k] - 1) * CS, CS, &cbuf[k * CS], [cs](size_t) { return cs; });
        if (rc) return rc;
    }
    launch_init_fields(m->dm, m->cur, m->d_lists, (int)nl, m->ph, 0.0, m->stream, nullptr, 0, m->d_iR);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(m->stream));
    return AMRLBM_OK;
}

// AMR_LBM_Rising_Bubble (AMR_Rising_2D_OMP.f90:51-67) for every instance at
// once. Initial_Mesh refines, level by level and in id order, the blocks whose
// instance's min_levels lies above their level; pass r of the init-regrid
// loop regrids the instances with r < max_levels - min_levels (the others are
// frozen: a pass beyond an instance's own count could derefine it) and sets
// up the fields of every leaf again (a function of the mesh alone, so a frozen
// instance's fields come out the same).
static int setup_ensemble_mesh(amrlbm_mesh *m)
{
    const int n = m->ens;
    reset_topology(m);
    m->cur = 0;
    for (int N = 1; N <= 2 * n; N++) {   // Initialize_Base_Block (AMR_Rising_2D_OMP.f90:79-115)
        const amrlbm_instance &q = m->inst[(N - 1) / 2];
        Block &b = m->blk[N];
        b.level = 0;
        b.dx = q.L0 / NX;
        b.I = 0;
        b.J = N - 1;
        b.has_child = true;
        for (int k = 0; k < 4; k++) b.child[k] = -1;
        b.parent = -1;
        b.isleaf = true;
        m->nblks = N;
        place(m, N);
    }
    const size_t S = (size_t)m->prm.max_blocks;
    HIPCHK(hipMemsetAsync(m->dm.pop[0], 0, esz(m) * POPB * S, m->stream));
    HIPCHK(hipMemsetAsync(m->dm.C, 0, esz(m) * CS * S, m->stream));
    HIPCHK(hipMemsetAsync(m->dm.PU, 0, esz(m) * PUB * S, m->stream));
    int top = 0, passes = 0;
    for (const amrlbm_instance &q : m->inst) {
        top = std::max(top, q.min_levels);
        passes = std::max(passes, q.max_levels - q.min_levels);
    }
    for (int L = 1; L <= top; L++) {   // Initial_Mesh (AMR_Rising_2D_OMP.f90:493-513)
        const int nend = m->nblks;
        for (int id = 1; id <= nend; id++) {
            const Block &b = m->blk[id];
            if (b.level == L - 1 && b.isleaf && L <= min_level_of(m, b)) {
                int rc = add_child_blocks(m, id, m->nblks + 1);
                if (rc) return rc;
            }
        }
    }
    array_of_leaf_blocks(m);
    determine_min_max_levels(m);
    int rc = init_ensemble(m);
    if (rc) return rc;
    std::vector<char> frozen(n, 0);
    for (int r = 0; r < passes; r++) {
        for (int q = 0; q < n; q++) frozen[q] = r >= m->inst[q].max_levels - m->inst[q].min_levels;
        determine_min_max_levels(m);
        if ((rc = regrid(m, 0.002, 0.001, nullptr, &frozen))) return rc;
        if ((rc = init_ensemble(m))) return rc;
    }
    m->T = 0;
    determine_min_max_levels(m);
    m->ready = true;
    return check_err_flags(m);
}

extern "C" {

int amrlbm_setup_ensemble(amrlbm_mesh *m, int n, const amrlbm_instance *inst)
{
    if (!m || n < 1 || !inst) return fail(AMRLBM_ERR_ARG, "bad ensemble arguments");
    if (m->prm.nroots != 2 * n) return fail(AMRLBM_ERR_ARG, "an ensemble of n instances needs nroots = 2 n");
    for (int q = 0; q < n; q++) {
        const amrlbm_instance &d = inst[q];
        if (d.min_levels < 0 || d.max_levels < d.min_levels || d.max_levels > m->prm.max_levels ||
            d.L0 != (NX << d.max_levels))
            return fail(AMRLBM_ERR_ARG, "instance " + std::to_string(q) +
                                            ": bad level range (L0 = nx 2^max_levels <= the mesh's L0)");
        if (!(d.R > 0.0) || !(d.W > 0.0) || !(d.tauh > 0.0) || !(d.taul > 0.0) || !(d.rhoh > d.rhol) ||
            !(d.rhol > 0.0))
            return fail(AMRLBM_ERR_ARG, "instance " + std::to_string(q) + ": bad fluid or bubble");
    }
    clear_ensemble(m);
    if (m->ens_cap < n) {
        (void)hipFree(m->d_phs);
        (void)hipFree(m->d_phsf);
        (void)hipFree(m->d_lv);
        (void)hipFree(m->d_iR);
        (void)hipFree(m->d_diag_inst);
        m->d_phs = nullptr;
        m->d_phsf = nullptr;
        m->d_lv = nullptr;
        m->d_iR = nullptr;
        m->d_diag_inst = nullptr;
        m->ens_cap = 0;
        HIPCHK(hipMalloc(&m->d_phs, sizeof(Phys) * n));
        HIPCHK(hipMalloc(&m->d_phsf, sizeof(PhysF) * n));
        HIPCHK(hipMalloc(&m->d_lv, sizeof(int) * 2 * n));
        HIPCHK(hipMalloc(&m->d_iR, sizeof(double) * n));
        HIPCHK(hipMalloc(&m->d_diag_inst, sizeof(double) * 9 * n));
        m->ens_cap = n;
    }
    std::vector<Phys> ph(n);
    std::vector<PhysF> phf(n);
    std::vector<int> lv(2 * n);
    std::vector<double> R(n);
    for (int q = 0; q < n; q++) {
        const amrlbm_instance &d = inst[q];
        amrlbm_params p = m->prm;   // the instance's constants as a one-domain mesh would hold them
        p.M0 = d.M0; p.W = d.W; p.sigma = d.sigma; p.beta = d.beta; p.kappa = d.kappa;
        p.taul = d.taul; p.tauh = d.tauh; p.rhol = d.rhol; p.rhoh = d.rhoh; p.gy = d.gy;
        ph[q] = phys_from(&p);
        phf[q] = physf_from(&p);
        lv[2 * q] = d.min_levels;
        lv[2 * q + 1] = d.max_levels;
        R[q] = d.R;
    }
    HIPCHK(hipMemcpyAsync(m->d_phs, ph.data(), sizeof(Phys) * n, hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(m->d_phsf, phf.data(), sizeof(PhysF) * n, hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(m->d_lv, lv.data(), sizeof(int) * 2 * n, hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(m->d_iR, R.data(), sizeof(double) * n, hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    m->ens = n;
    m->inst.assign(inst, inst + n);
    m->tile_roots = 2;
    m->dm.tile_roots = 2;
    m->dm.phs = m->d_phs;
    m->dm.phsf = m->d_phsf;
    m->dm.lv = m->d_lv;
    int rc = setup_ensemble_mesh(m);
    if (rc) clear_ensemble(m);
    return rc;
}

int amrlbm_output_instances(amrlbm_mesh *m, int n, double *out)
{
    NEED_READY(m);
    if (!out || !m->ens || n != m->ens) return fail(AMRLBM_ERR_ARG, "not an ensemble mesh of n instances");
    int rc = upload_tables(m);
    if (rc) return rc;
    launch_diag(m->dm, m->d_lists, (int)m->leaf.size(), m->ph, m->d_part, m->d_diag, m->stream, n,
                m->d_diag_inst);
    HIPCHK(hipGetLastError());
    std::vector<double> d(9 * (size_t)n);
    HIPCHK(hipMemcpyAsync(d.data(), m->d_diag_inst, sizeof(double) * d.size(), hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    double bad = 0.0;
    for (int q = 0; q < n; q++) {
        const double *a = &d[9 * (size_t)q];
        double *o = out + 9 * (size_t)q;
        for (int k = 0; k < 9; k++) o[k] = a[k];
        o[5] = a[5] / a[4];
        o[6] = a[6] / a[4];
        o[7] = a[7] / a[4];
        bad += a[8];
    }
    if (bad != 0.0) return fail(AMRLBM_ERR_NONFINITE, "non-finite phase field in " + std::to_string((long)bad) + " cells");
    return AMRLBM_OK;
}

int amrlbm_get_counts(amrlbm_mesh *m, int counts[7])
{
    if (!m || !counts) return fail(AMRLBM_ERR_ARG, "null argument");
    counts[0] = m->nblks;
    counts[1] = (int)m->leaf.size();
    counts[2] = m->levelmin;
    counts[3] = m->levelmax;
    counts[4] = m->kk;
    counts[5] = (int)m->freeq.size();
    counts[6] = m->T;
    return AMRLBM_OK;
}

int amrlbm_set_step(amrlbm_mesh *m, int T)
{
    if (!m || T < 0) return fail(AMRLBM_ERR_ARG, "bad step counter");
    m->T = T;
    return AMRLBM_OK;
}

int amrlbm_get_topology(amrlbm_mesh *m, int *blocks, int *freeq)
{
    if (!m || !blocks) return fail(AMRLBM_ERR_ARG, "null argument");
    for (int n = 1; n <= m->nblks; n++) {
        const Block &b = m->blk[n];
        int *r = blocks + (size_t)(n - 1) * 21;
        r[0] = b.level; r[1] = b.dx; r[2] = b.I; r[3] = b.J; r[4] = b.parent;
        r[5] = b.isleaf; r[6] = b.hasleaf; r[7] = b.refin; r[8] = b.deref;
        for (int k = 0; k < 4; k++) r[9 + k] = (b.level >= 0 && b.has_child) ? b.child[k] : 0;
        for (int k = 1; k <= 8; k++) r[12 + k] = b.level >= 0 ? nb_value(m, n, k) : 0;
    }
    if (freeq) {
        int k = 0;
        for (int v : m->freeq) freeq[k++] = v;
    }
    return AMRLBM_OK;
}

static int set_topology(amrlbm_mesh *m, int nblks, const int *blocks, int kk, int nfree, const int *freeq, int T,
                        int ninst)
{
    if (!m || !blocks || nblks < 1 || nblks > m->prm.max_blocks || nfree < 0 || (nfree > 0 && !freeq))
        return fail(AMRLBM_ERR_ARG, "bad topology arguments");
    if (ninst < 1 || m->prm.nroots % ninst) return fail(AMRLBM_ERR_ARG, "the roots must split evenly into the instances");
    clear_ensemble(m);
    m->tile_roots = m->prm.nroots / ninst;
    reset_topology(m);
    m->nblks = nblks;
    // a rejected table leaves no mesh behind (every later call reports it)
    auto bad = [&](int n, const std::string &what) {
        reset_topology(m);
        m->ready = false;
        return fail(AMRLBM_ERR_TOPOLOGY, "block " + std::to_string(n) + ": " + what);
    };
    for (int n = 1; n <= nblks; n++) {
        const int *r = blocks + (size_t)(n - 1) * 21;
        Block &b = m->blk[n];
        b.level = r[0]; b.dx = r[1]; b.I = r[2]; b.J = r[3]; b.parent = r[4];
        b.isleaf = r[5]; b.hasleaf = r[6]; b.refin = r[7]; b.deref = r[8];
        b.has_child = b.level >= 0;
        b.alloc = b.level >= 0 && b.isleaf;   // load_snapshot sets it from the file
        for (int k = 0; k < 4; k++) b.child[k] = r[9 + k];
        if (b.level > m->prm.max_levels) return bad(n, "level above max_levels");
        if (b.level >= 0) {
            if (b.I < 0 || b.I >= m->width(b.level) || b.J < 0 || b.J >= m->height(b.level))
                return bad(n, "block position outside the domain");
            if (b.dx != (m->prm.L0 / NX) >> b.level) return bad(n, "dx is not L0 / nx / 2^Level");
            place(m, n);
        }
    }
    for (int n = 1; n <= nblks; n++) {
        const Block &b = m->blk[n];
        if (b.level < -1) return bad(n, "level below -1");
        if (b.level < 0) continue;
        // a root's Parent is -1 (AMR_Rising_2D_OMP.f90:95)
        if (b.level == 0 ? b.parent != -1 : (b.parent < 1 || b.parent > nblks))
            return bad(n, "parent id not -1 for a root or in 1..Nblks below");
        if (b.level > 0) {
            const Block &p = m->blk[b.parent];
            if (p.level != b.level - 1 || p.isleaf || n < p.child[0] || n > p.child[0] + 3)
                return bad(n, "parent " + std::to_string(b.parent) + " does not list it as a child");
        }
        if (!b.isleaf) {
            if (b.child[0] < 1 || b.child[0] > nblks - 3) return bad(n, "first child id outside 1..Nblks-3");
            for (int k = 0; k < 4; k++) {
                const Block &c = m->blk[b.child[0] + k];
                if (b.child[k] != b.child[0] + k || c.level != b.level + 1 || c.parent != n)
                    return bad(n, "children are not 4 contiguous blocks one level finer");
                if (b.hasleaf && !c.isleaf) return bad(n, "hasLeaf set with a refined child");
            }
        } else if (b.hasleaf) {
            return bad(n, "hasLeaf set on a leaf");
        }
    }
    if (kk != 4 * nfree) return bad(0, "KK is not 4 times the free-list length");
    for (int k = 0; k < nfree; k++) {
        const int f = freeq[k];
        if (f < 1 || f > nblks - 3) return bad(f, "free-list id outside 1..Nblks-3");
        for (int c = f; c < f + 4; c++)
            if (m->blk[c].level != -1) return bad(c, "on the free list but in use");
    }
    for (int n = 1; n <= nblks; n++) {
        if (m->blk[n].level < 0) continue;
        for (int k = 1; k <= 8; k++)
            if (nb_value(m, n, k) != blocks[(size_t)(n - 1) * 21 + 12 + k])
                return bad(n, "neighbour " + std::to_string(k) + " disagrees with the quadtree");
    }
    for (int k = 0; k < nfree; k++) m->freeq.push_back(freeq[k]);
    m->kk = kk;
    m->T = T;
    array_of_leaf_blocks(m);
    determine_min_max_levels(m);
    m->ready = true;
    return upload_tables(m);
}

int amrlbm_set_topology(amrlbm_mesh *m, int nblks, const int *blocks, int kk, int nfree, const int *freeq, int T)
{
    return set_topology(m, nblks, blocks, kk, nfree, freeq, T, 1);   // a caller's mesh is one instance
}

// Snapshot file: the layout of the reference dump harness and the oracle
// (oracle/ref_harness.f90, oracle/amroracle.c): 16-int header, the free list,
// then per block 22 ints (the topology row and a has-arrays flag) and, for
// leaves, h, g, C, P, Ux, Uy in the reference's array layouts, FP64.
// Header word 15 (unused there) holds the instance count of a replicated mesh.
constexpr int kSnapMagic = 1095586380;
static int put_block(amrlbm_mesh *m, int id, const double *h, const double *g, const double *C, const double *P,
                     const double *Ux, const double *Uy, bool shifted);
static int get_block(amrlbm_mesh *m, int id, double *h, double *g, double *C, double *P, double *Ux, double *Uy,
                     bool shifted);

int amrlbm_save_snapshot(amrlbm_mesh *m, const char *path)
{
    if (!m || !path) return fail(AMRLBM_ERR_ARG, "null argument");
    if (!m->ready) return fail(AMRLBM_ERR_STATE, "mesh not set up");
    if (m->ens) return fail(AMRLBM_ERR_ARG, "an ensemble mesh has no restart file (one fluid per file)");
    HIPCHK(hipStreamSynchronize(m->stream));
    std::vector<int> rows((size_t)m->nblks * 21), freeq(m->freeq.size() + 1);
    int rc = amrlbm_get_topology(m, rows.data(), freeq.data());
    if (rc) return rc;
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return fail(AMRLBM_ERR_ARG, std::string("cannot write ") + path);
    const amrlbm_params &p = m->prm;
    const bool dev = m->dm.f32 != 0;   // an FP32 mesh writes its deviations, exactly
    int hdr[16] = {kSnapMagic, dev ? 2 : 1, p.nx, p.ny, p.L0, p.min_levels, p.max_levels, m->nblks, (int)m->leaf.size(), m->T,
                   m->levelmin, m->levelmax, m->kk, (int)m->freeq.size(), p.max_blocks,
                   m->tile_roots < p.nroots ? p.nroots / m->tile_roots : 0};
    bool ok = std::fwrite(hdr, sizeof(int), 16, fp) == 16;
    ok = ok && std::fwrite(freeq.data(), sizeof(int), m->freeq.size(), fp) == m->freeq.size();
    const size_t nf = 9 * FS, nc = CS, np = PS;
    std::vector<double> a(2 * nf + nc + 3 * np);
    for (int n = 1; n <= m->nblks && ok; n++) {
        int iv[22];
        std::memcpy(iv, &rows[(size_t)(n - 1) * 21], sizeof(int) * 21);
        const Block &b = m->blk[n];
        iv[21] = (b.level >= 0 && (b.isleaf || b.alloc)) ? 1 : 0;
        ok = std::fwrite(iv, sizeof(int), 22, fp) == 22;
        if (!ok || !iv[21]) continue;
        double *d = a.data();
        rc = get_block(m, n, d, d + nf, d + 2 * nf, d + 2 * nf + nc, d + 2 * nf + nc + np, d + 2 * nf + nc + 2 * np,
                       !dev);
        if (rc) {
            std::fclose(fp);
            return rc;
        }
        ok = std::fwrite(d, sizeof(double), a.size(), fp) == a.size();
    }
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? AMRLBM_OK : fail(AMRLBM_ERR_ARG, std::string("short write to ") + path);
}

int amrlbm_load_snapshot(amrlbm_mesh *m, const char *path)
{
    if (!m || !path) return fail(AMRLBM_ERR_ARG, "null argument");
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return fail(AMRLBM_ERR_ARG, std::string("cannot read ") + path);
    auto bad = [&](const std::string &why) {
        std::fclose(fp);
        return fail(AMRLBM_ERR_ARG, std::string(path) + ": " + why);
    };
    int hdr[16];
    if (std::fread(hdr, sizeof(int), 16, fp) != 16 || hdr[0] != kSnapMagic || hdr[1] < 1 || hdr[1] > 2)
        return bad("not a snapshot");
    const bool file_dev = hdr[1] == 2;   // values are deviations: h - w_q, C - 1
    const amrlbm_params &p = m->prm;
    if (hdr[2] != p.nx || hdr[3] != p.ny || hdr[4] != p.L0 || hdr[5] != p.min_levels || hdr[6] != p.max_levels)
        return bad("mesh parameters differ from this mesh's");
    const int nblks = hdr[7], nfree = hdr[13], ninst = hdr[15] > 1 ? hdr[15] : 1;
    if (nblks < 1 || nblks > p.max_blocks || nfree < 0 || nfree > p.max_blocks) return bad("block counts out of range");
    std::vector<int> freeq(nfree + 1), rows((size_t)nblks * 21), leafs;
    if ((int)std::fread(freeq.data(), sizeof(int), nfree, fp) != nfree) return bad("truncated free list");
    const size_t nf = 9 * FS, nc = CS, np = PS, na = 2 * nf + nc + 3 * np;
    std::vector<double> arrays;
    for (int n = 1; n <= nblks; n++) {
        int iv[22];
        if (std::fread(iv, sizeof(int), 22, fp) != 22) return bad("truncated block table");
        std::memcpy(&rows[(size_t)(n - 1) * 21], iv, sizeof(int) * 21);
        if (!iv[21]) continue;
        leafs.push_back(n);
        arrays.resize(arrays.size() + na);
        if (std::fread(&arrays[arrays.size() - na], sizeof(double), na, fp) != na) return bad("truncated arrays");
    }
    std::fclose(fp);
    int rc = set_topology(m, nblks, rows.data(), hdr[12], nfree, freeq.data(), hdr[9], ninst);
    if (rc) return rc;
    for (int n : leafs) m->blk[n].alloc = true;
    const bool raw = file_dev && m->dm.f32;   // same storage: no conversion at all
    for (size_t k = 0; k < leafs.size(); k++) {
        double *d = &arrays[k * na];
        if (file_dev && !raw) {   // deviations into an FP64 mesh: absolute values
            for (size_t e = 0; e < nf; e++) d[e] += kW[e % 9];   // h(9, 0:nx+1, 0:ny+1): q fastest
            for (size_t e = 0; e < nc; e++) d[2 * nf + e] += 1.0;
        }
        rc = put_block(m, leafs[k], d, d + nf, d + 2 * nf, d + 2 * nf + nc, d + 2 * nf + nc + np,
                       d + 2 * nf + nc + 2 * np, !raw);
        if (rc) return rc;
    }
    return AMRLBM_OK;
}

// shifted: the host arrays hold absolute values (the FP32 mesh's storage
// shift is applied on the way); else they hold what the mesh stores
static int put_block(amrlbm_mesh *m, int id, const double *h, const double *g, const double *C, const double *P,
                     const double *Ux, const double *Uy, bool shifted)
{
    if (!m || id < 1 || id > m->prm.max_blocks) return fail(AMRLBM_ERR_ARG, "bad block id");
    const size_t s = (size_t)(id - 1);
    const bool f32 = m->dm.f32 && shifted;
    std::vector<double> buf(9 * GS);
    int rc = AMRLBM_OK;
    if (h) {
        ref_to_dev_pop(h, buf.data());
        rc = upload(m, m->dm.pop[m->cur], s * POPB, 9 * GS, buf.data(),
                    [f32](size_t k) { return f32 ? kW[k / GS] : 0.0; });
        if (rc) return rc;
    }
    if (g) {
        ref_to_dev_pop(g, buf.data());
        if ((rc = upload(m, m->dm.pop[m->cur], s * POPB + 9 * GS, 9 * GS, buf.data(), no_shift))) return rc;
    }
    if (C && (rc = upload(m, m->dm.C, s * CS, CS, C, [f32](size_t) { return f32 ? 1.0 : 0.0; }))) return rc;
    if (P && (rc = upload(m, m->dm.PU, s * PUB, PS, P, no_shift))) return rc;
    if (Ux && (rc = upload(m, m->dm.PU, s * PUB + PS, PS, Ux, no_shift))) return rc;
    if (Uy && (rc = upload(m, m->dm.PU, s * PUB + 2 * PS, PS, Uy, no_shift))) return rc;
    return AMRLBM_OK;
}

static int get_block(amrlbm_mesh *m, int id, double *h, double *g, double *C, double *P, double *Ux, double *Uy,
                     bool shifted)
{
    if (!m || id < 1 || id > m->prm.max_blocks) return fail(AMRLBM_ERR_ARG, "bad block id");
    const size_t s = (size_t)(id - 1);
    const bool f32 = m->dm.f32 && shifted;
    std::vector<double> buf(9 * GS);
    int rc = AMRLBM_OK;
    if (h) {
        rc = download(m, m->dm.pop[m->cur], s * POPB, 9 * GS, buf.data(),
                      [f32](size_t k) { return f32 ? kW[k / GS] : 0.0; });
        if (rc) return rc;
        dev_to_ref_pop(buf.data(), h);
    }
    if (g) {
        if ((rc = download(m, m->dm.pop[m->cur], s * POPB + 9 * GS, 9 * GS, buf.data(), no_shift))) return rc;
        dev_to_ref_pop(buf.data(), g);
    }
    if (C && (rc = download(m, m->dm.C, s * CS, CS, C, [f32](size_t) { return f32 ? 1.0 : 0.0; }))) return rc;
    if (P && (rc = download(m, m->dm.PU, s * PUB, PS, P, no_shift))) return rc;
    if (Ux && (rc = download(m, m->dm.PU, s * PUB + PS, PS, Ux, no_shift))) return rc;
    if (Uy && (rc = download(m, m->dm.PU, s * PUB + 2 * PS, PS, Uy, no_shift))) return rc;
    return AMRLBM_OK;
}

int amrlbm_put_block(amrlbm_mesh *m, int id, const double *h, const double *g, const double *C, const double *P,
                     const double *Ux, const double *Uy)
{
    return put_block(m, id, h, g, C, P, Ux, Uy, true);
}

}  // extern "C"

// ---- Tecplot output: Fortran edit descriptors as the reference's runtime
// writes them (each right-justified in its width, asterisks when too wide) --
namespace {

std::string fit(const std::string &t, int w)
{
    if ((int)t.size() > w) return std::string(w, '*');
    return std::string(w - t.size(), ' ') + t;
}

std::string fort_i(long long v, int w) { return fit(std::to_string(v), w); }

// Fw.d: the binary value rounded to d decimals (ties to even), leading zero kept
std::string fort_f(double x, int w, int d)
{
    char b[64];
    std::snprintf(b, sizeof b, "%.*f", d, x);
    return fit(b, w);
}

// Ew.d: 0.d1..dd E+ee, the exponent without the E when it needs three digits
std::string fort_e(double x, int w, int d)
{
    std::string t = std::signbit(x) ? "-" : "";
    if (x == 0.0) {
        t += "0." + std::string(d, '0') + "E+00";
        return fit(t, w);
    }
    char b[64];
    std::snprintf(b, sizeof b, "%.*e", d - 1, std::fabs(x));   // d.ddddde+XX, d significant digits
    const char *e = std::strchr(b, 'e');
    std::string digits(b, e - b);
    digits.erase(1, 1);   // the point
    const int ex = std::atoi(e + 1) + 1;
    char eb[16];
    if (ex >= -99 && ex <= 99)
        std::snprintf(eb, sizeof eb, "E%c%02d", ex < 0 ? '-' : '+', std::abs(ex));
    else
        std::snprintf(eb, sizeof eb, "%c%03d", ex < 0 ? '-' : '+', std::abs(ex));
    t += "0." + digits + eb;
    return fit(t, w);
}

// List-directed REAL(8): the fewest significant digits that read back to the
// same value (of the equally short forms, the truncation of the value where
// it reads back, else the rounded one), in F form without a leading zero when
// the value rounded to one digit is 0.d E+e with 0 <= e <= 15, else d.ddE+ee.
std::string fort_list_real(double x)
{
    std::string sign = std::signbit(x) ? "-" : "";
    if (x == 0.0) return sign + "0.";
    const double a = std::fabs(x);
    char b[96];
    int L = 1;
    for (; L < 17; L++) {
        std::snprintf(b, sizeof b, "%.*e", L - 1, a);
        if (std::strtod(b, nullptr) == a) break;
    }
    std::snprintf(b, sizeof b, "%.*e", L - 1, a);   // rounded, L digits
    char ex40[96];
    std::snprintf(ex40, sizeof ex40, "%.40e", a);
    const char *e40 = std::strchr(ex40, 'e');
    std::string all(ex40, e40 - ex40);
    all.erase(1, 1);
    const std::string trunc = all.substr(0, L);
    std::string tstr = trunc.substr(0, 1) + "." + trunc.substr(1) + e40;
    std::string D;
    int E10;
    if (std::strtod(tstr.c_str(), nullptr) == a) {
        D = trunc;
        E10 = std::atoi(e40 + 1);
    } else {
        const char *e = std::strchr(b, 'e');
        D = std::string(b, e - b);
        D.erase(1, 1);
        E10 = std::atoi(e + 1);
    }
    while (D.size() > 1 && D.back() == '0') D.pop_back();
    char one[32];
    std::snprintf(one, sizeof one, "%.0e", a);
    const int expo1 = std::atoi(std::strchr(one, 'e') + 1) + 1;
    std::string t;
    if (expo1 >= 0 && expo1 <= 15) {
        const int p = E10 + 1;
        if (p <= 0)
            t = "." + std::string(-p, '0') + D;
        else if (p >= (int)D.size())
            t = D + std::string(p - D.size(), '0') + ".";
        else
            t = D.substr(0, p) + "." + D.substr(p);
    } else {
        char eb[16];
        std::snprintf(eb, sizeof eb, "E%c%02d", E10 < 0 ? '-' : '+', std::abs(E10));
        t = D.substr(0, 1) + "." + D.substr(1) + eb;
    }
    return sign + t;
}

// records of `per` values each, as a format with a repeat count writes them
template <class V, class F>
void records(std::string &out, const std::vector<V> &v, int per, F fmt)
{
    for (size_t k = 0; k < v.size(); k++) {
        out += fmt(v[k]);
        if ((k + 1) % per == 0 || k + 1 == v.size()) out += '\n';
    }
}

}  // namespace

extern "C" {

int amrlbm_tecplot_cell_center(amrlbm_mesh *m, const char *path, double D)
{
    if (!m) return fail(AMRLBM_ERR_ARG, "null mesh");
    if (!m->ready) return fail(AMRLBM_ERR_STATE, "mesh not set up");
    if (!path || !(D > 0.0)) return fail(AMRLBM_ERR_ARG, "bad Tecplot arguments");
    if (m->ens) return fail(AMRLBM_ERR_ARG, "an ensemble mesh has no Tecplot file (one domain per file)");
    const size_t nb = (size_t)m->nblks;
    std::vector<double> C(nb * CS), PU(nb * PUB);
    const bool f32 = m->dm.f32;
    int rc = download(m, m->dm.C, 0, nb * CS, C.data(), [f32](size_t) { return f32 ? 1.0 : 0.0; });
    if (rc) return rc;
    if ((rc = download(m, m->dm.PU, 0, nb * PUB, PU.data(), no_shift))) return rc;
    // TecPlot_CellCenter (TecPlot_ASCII.f90:21-77): leaves in Array_of_Leaf_Blocks order
    const long long nl = (long long)m->leaf.size();
    const long long Nn = nl * (NX + 1) * (NY + 1), Ne = nl * NX * NY;
    std::vector<float> X, Y;
    std::vector<double> u, v, c, p;
    std::vector<long long> conn;
    X.reserve(Nn); Y.reserve(Nn); c.reserve(Nn);
    u.reserve(Ne); v.reserve(Ne); p.reserve(Ne);
    conn.reserve(4 * Ne);
    const float L0 = (float)m->prm.L0;
    const double sigma = m->ph.sigma;
    long long N = 1;
    for (int b : m->leaf) {
        const Block &B = m->blk[b];
        const int dx = B.dx;
        const double *Cb = C.data() + (size_t)(b - 1) * CS, *PUb = PU.data() + (size_t)(b - 1) * PUB;
        auto Cat = [Cb](int x, int y) { return Cb[(y + 1) * CW + (x + 1)]; };
        for (int J = 1; J <= NY + 1; J++)
            for (int I = 1; I <= NX + 1; I++) {
                const int xn = B.I * dx * NX + dx * (I - 1), yn = B.J * dx * NY + dx * (J - 1);
                X.push_back((float)xn / L0);   // default-kind REAL division
                Y.push_back((float)yn / L0);
                c.push_back(0.25 * (((Cat(I - 1, J - 1) + Cat(I, J - 1)) + Cat(I - 1, J)) + Cat(I, J)) - 0.5);
                if (I <= NX && J <= NY) {
                    const int k = (I - 1) + NX * (J - 1);
                    u.push_back(PUb[PS + k]);
                    v.push_back(PUb[2 * PS + k]);
                    p.push_back(PUb[k] * 0.5 * D / sigma);
                    conn.insert(conn.end(), {N, N + 1, N + 1 + NX + 1, N + NX + 1});
                }
                N++;
            }
    }
    std::string out;
    out.reserve((size_t)(Nn * 34 + Ne * 70 + 256));
    out += " Variables = X, Y, Ux, Uy, C, P\n";
    out += "Zone N =" + fort_i(Nn, 6) + ", E =" + fort_i(Ne, 6) + ", ZONETYPE=FEquadrilateral, DATAPACKING=Block\n";
    out += " VARLOCATION=([3,4,6]=CELLCENTERED), T = \" " + fort_list_real(m->T * std::sqrt(m->ph.gy / D)) + " \"\n";
    records(out, X, 9, [](float q) { return fort_f(q, 10, 3); });
    records(out, Y, 9, [](float q) { return fort_f(q, 10, 3); });
    records(out, u, 8, [](double q) { return fort_e(q, 14, 6); });
    records(out, v, 8, [](double q) { return fort_e(q, 14, 6); });
    records(out, c, 9, [](double q) { return fort_e(q, 14, 6); });
    records(out, p, 8, [](double q) { return fort_e(q, 14, 6); });
    records(out, conn, 8, [](long long q) { return fort_i(q, 7); });
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(AMRLBM_ERR_ARG, std::string("cannot open ") + path);
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    if (std::fclose(f) != 0 || !ok) return fail(AMRLBM_ERR_ARG, std::string("cannot write ") + path);
    return AMRLBM_OK;
}

int amrlbm_get_block(amrlbm_mesh *m, int id, double *h, double *g, double *C, double *P, double *Ux, double *Uy)
{
    return get_block(m, id, h, g, C, P, Ux, Uy, true);
}

int amrlbm_determine_min_max_levels(amrlbm_mesh *m, int *levelmin, int *levelmax)
{
    if (!m) return fail(AMRLBM_ERR_ARG, "null mesh");
    determine_min_max_levels(m);
    if (levelmin) *levelmin = m->levelmin;
    if (levelmax) *levelmax = m->levelmax;
    return AMRLBM_OK;
}

int amrlbm_fill_ghost_cells_c2(amrlbm_mesh *m)
{
    NEED_READY(m);
    int rc = phase_fill_c(m, KIND_C2, OP_NONE);
    return rc ? rc : check_err_flags(m);
}

int amrlbm_collide(amrlbm_mesh *m)
{
    NEED_READY(m);
    return phase_collide(m);
}

int amrlbm_fill_ghost_cells_hg(amrlbm_mesh *m)
{
    NEED_READY(m);
    int rc = phase_fill_hg(m);
    return rc ? rc : check_err_flags(m);
}

int amrlbm_stream(amrlbm_mesh *m)
{
    NEED_READY(m);
    return phase_stream(m, false);
}

int amrlbm_phase_field_lbm_h_g(amrlbm_mesh *m, double ms[3])
{
    NEED_READY(m);
    float t[3] = {0, 0, 0};
    hipEvent_t e[4];
    for (int k = 0; k < 4; k++) HIPCHK(hipEventCreate(&e[k]));
    HIPCHK(hipEventRecord(e[0], m->stream));
    int rc = phase_collide(m);
    if (!rc) {
        HIPCHK(hipEventRecord(e[1], m->stream));
        rc = phase_fill_hg(m);
    }
    if (!rc) {
        HIPCHK(hipEventRecord(e[2], m->stream));
        rc = phase_stream(m, false);
    }
    if (!rc) {
        HIPCHK(hipEventRecord(e[3], m->stream));
        HIPCHK(hipEventSynchronize(e[3]));
        for (int k = 0; k < 3; k++) HIPCHK(hipEventElapsedTime(&t[k], e[k], e[k + 1]));
        if (ms)
            for (int k = 0; k < 3; k++) ms[k] = t[k];
    }
    for (int k = 0; k < 4; k++) (void)hipEventDestroy(e[k]);
    return rc ? rc : check_err_flags(m);
}

int amrlbm_macroscopic_values_h(amrlbm_mesh *m)
{
    NEED_READY(m);
    return phase_macro_h(m);
}

int amrlbm_fill_ghost_cells_c(amrlbm_mesh *m)
{
    NEED_READY(m);
    int rc = phase_fill_c(m, KIND_C1, OP_NONE);
    return rc ? rc : check_err_flags(m);
}

int amrlbm_macroscopic_values_g(amrlbm_mesh *m)
{
    NEED_READY(m);
    return phase_macro_g(m);
}

int amrlbm_refine_derefine_mesh(amrlbm_mesh *m, double refin_tol, double deref_tol)
{
    NEED_READY(m);
    return regrid(m, refin_tol, deref_tol);
}

int amrlbm_regrid_check(amrlbm_mesh *m, int out[2])
{
    NEED_READY(m);
    if (!out) return fail(AMRLBM_ERR_ARG, "null argument");
    int rc = upload_tables(m);
    if (rc) return rc;
    determine_min_max_levels(m);
    // C, its first ghost ring and U, P are current between steps
    launch_err(m->dm, all_leaves(m), nleaves(m), m->d_err, m->d_err_slot, m->stream);
    HIPCHK(launch_regrid_check(m->dm, m->nblks, m->d_err_slot, m->prm.min_levels, m->prm.max_levels, m->levelmin,
                               m->levelmax, 0.002, 0.001, m->d_changed, m->stream));
    HIPCHK(hipMemcpyAsync(m->h_changed, m->d_changed, sizeof(int) * 2, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    out[0] = m->h_changed[0];
    out[1] = m->h_changed[1];
    HIPCHK(hipMemsetAsync(m->d_changed, 0, sizeof(int) * 2, m->stream));   // the skip word stays clear
    HIPCHK(hipStreamSynchronize(m->stream));
    return AMRLBM_OK;
}

int amrlbm_refine_derefine_predicted(amrlbm_mesh *m, const int *ids, int n)
{
    NEED_READY(m);
    if (n < 0 || (n > 0 && !ids)) return fail(AMRLBM_ERR_ARG, "bad predicted blocks");
    std::vector<char> extra(m->nblks + 1, 0);
    for (int k = 0; k < n; k++)
        if (ids[k] >= 1 && ids[k] <= m->nblks) extra[ids[k]] = 1;
    return regrid(m, 0.002, 0.001, &extra);
}

}  // extern "C"

// The end of a stepping call (ev0 recorded at its start). tail_checked: the
// call's last stream work is a decision copy the host has waited for
// (ev_check); that copy carried the error word, and its event ends the
// call's timing, so no further round trip is needed. Otherwise the error
// word is copied behind the call's work and one wait (ev1) covers both.
// The segment kernel's barrier word is read (check_err_flags) only after a
// call that ran it.
static int end_of_steps(amrlbm_mesh *m, bool tail_checked)
{
    if (m->seg_kernel_ran) {
        HIPCHK(hipEventRecord(m->ev1, m->stream));
        HIPCHK(hipEventSynchronize(m->ev1));
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, m->ev0, m->ev1));
        m->last_step_ms = ms;
        return check_err_flags(m);
    }
    hipEvent_t end = m->ev_check;
    if (!tail_checked) {
        HIPCHK(hipMemcpyAsync(m->h_changed + 3, m->d_err_flags, sizeof(int), hipMemcpyDeviceToHost, m->stream));
        HIPCHK(hipEventRecord(m->ev1, m->stream));
        HIPCHK(hipEventSynchronize(m->ev1));
        end = m->ev1;
    }
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, m->ev0, end));
    m->last_step_ms = ms;
    if (m->h_changed[3]) {
        HIPCHK(hipMemsetAsync(m->d_err_flags, 0, sizeof(int), m->stream));
        return fail(AMRLBM_ERR_TOPOLOGY, "ghost fill met a neighbour more than one level apart");
    }
    return AMRLBM_OK;
}

// The steps of amrlbm_step. Stream order: segment r, regrid check r, segment
// r+1, ... The host reads check r while segment r+1 runs. When check r says the
// mesh changes, segment r+1 has skipped all its work: the host takes its steps
// back, runs Refine_Derefine_Mesh on the state of check r and goes on from
// there. *ahead_len is set while a segment launched behind a pending check
// has advanced T (so an error return can take it back).
static int step_segments(amrlbm_mesh *m, int nsteps, int regrid_every, int *ahead_len)
{
    int rc = AMRLBM_OK;
    HIPCHK(hipEventRecord(m->ev0, m->stream));
    m->seg_kernel_ran = false;
    int done = 0;
    bool pending = false, tail_checked = false;
    while (done < nsteps || pending) {
        bool launched = false;
        int len = 0;
        tail_checked = false;
        if (done < nsteps) {
            len = segment_length(m, nsteps - done, regrid_every);
            if (!pending) {
                determine_min_max_levels(m);
                if ((rc = upload_tables(m))) return rc;
            }
            if ((rc = launch_steps(m, len))) return rc;
            m->T += len;
            done += len;
            launched = true;
            if (pending) *ahead_len = len;
        }
        if (pending) {
            pending = false;
            HIPCHK(hipEventSynchronize(m->ev_check));
            if (*m->h_changed) {
                // with the device's flags the host regrids at once: the
                // segment launched behind the decision drains meanwhile (its
                // kernels skip), and every change to the device state is
                // queued behind it; without them the host flags from a fresh
                // Err_Interface once the stream is idle
                const bool dev_flags = m->h_changed[4] == 1;
                if (!dev_flags) HIPCHK(hipStreamSynchronize(m->stream));
                m->device_flag_regrids += dev_flags;
                if (launched) {
                    m->T -= len;
                    done -= len;
                    m->segments_skipped++;
                }
                *ahead_len = 0;
                m->regrid_changes++;
                // the skip word stays set while the host regrids (no step
                // kernel runs meanwhile); it is cleared below, or by the
                // caller on an error return
                if ((rc = regrid(m, 0.002, 0.001, nullptr, nullptr, nullptr, nullptr, dev_flags ? m->h_changed + 4 : nullptr)))
                    return rc;
                HIPCHK(hipMemsetAsync(m->d_changed, 0, sizeof(int), m->stream));
                continue;
            }
            *ahead_len = 0;
            clear_regrid_flags(m);
            tail_checked = !launched;
        }
        if (launched && m->T % regrid_every == 0) {
            if ((rc = enqueue_regrid_check(m))) return rc;
            pending = true;
        }
    }
    determine_min_max_levels(m);
    return end_of_steps(m, tail_checked);
}

// The learned predictor's refinement flags of every leaf at the current state,
// by block slot in d_pred (network inputs, forward pass and flags, on the
// device), for the regrid that follows. fresh_err: d_err_slot already holds
// Err_Interface of this state (a segment's end kernel ran on it).
static int enqueue_prediction(amrlbm_mesh *m, bool fresh_err)
{
    const int n = nleaves(m);
    if (n > m->net_cap) {
        const int cap = n + n / 2 + 64;
        (void)hipFree(m->d_net_x);
        (void)hipFree(m->d_net_logit);
        m->d_net_x = m->d_net_logit = nullptr;
        HIPCHK(hipMalloc(&m->d_net_x, sizeof(float) * 3 * 18 * 18 * (size_t)cap));
        HIPCHK(hipMalloc(&m->d_net_logit, sizeof(float) * 2 * (size_t)cap));
        m->net_cap = cap;
    }
    if (!fresh_err) launch_err(m->dm, all_leaves(m), n, m->d_err, m->d_err_slot, m->stream);
    launch_cnn_features(m->dm, all_leaves(m), n, m->d_net_x, m->ph.W, m->stream);
    launch_cnn(m->d_net_x, nullptr, n, m->d_net_w, m->d_net_logit, nullptr, 0.f, nullptr, nullptr, m->stream);
    // a prediction counts where the interface's profile already reaches the
    // block (Err_Interface above net_gate times the refinement tolerance)
    launch_pred_flags(m->d_net_logit, all_leaves(m), n, m->net_threshold, m->d_err_slot, m->net_gate * 0.002,
                      m->d_pred, m->stream);
    HIPCHK(hipGetLastError());
    return AMRLBM_OK;
}

// The decision after a judged segment of len steps that ends at state T: at a
// full regrid (T a multiple of the full-regrid interval) the plain criterion
// over every leaf, unless the judge stopped the segment before its last
// state; between full regrids, the judge alone. Its copy (decision, any
// refinement, stop ordinal) reaches h_changed when ev_check fires.
static int enqueue_learned_check(amrlbm_mesh *m, int len, bool full)
{
    if (full)
        HIPCHK(launch_regrid_check(m->dm, m->nblks, m->d_err_slot, m->prm.min_levels, m->prm.max_levels, m->levelmin,
                                   m->levelmax, 0.002, 0.001, m->d_changed, m->stream, m->d_judge, len));
    else
        launch_judge_check(m->d_judge, len + 1, m->d_changed, m->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(m->h_changed, m->d_changed, sizeof(int) * 4, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipEventRecord(m->ev_check, m->stream));
    m->regrid_checks++;
    return AMRLBM_OK;
}

// The learned regrid's cadence (no counterpart in the reference). Full
// regrids every full_every steps take the plain criterion; between them the
// plain criterion judges every state on the device, and the first state it
// finds needing refinement is regridded at once (a miss). The network never
// changes the mesh on its own: its flags join every regrid that does happen,
// so that blocks the interface is about to reach are refined ahead and the
// misses they would cause do not come. Stream order as in amrlbm_step:
// segment r, decision r, segment r+1, ...; the host reads decision r while
// segment r+1 runs, and a segment behind a "changed" decision skips all its
// work and is taken back. *ahead_len as in step_segments.
static int step_learned(amrlbm_mesh *m, int nsteps, int full_every, long long st[4], int *ahead_len)
{
    int rc = AMRLBM_OK, done = 0, plen = 0;
    bool pending = false, pfull = false, tail_checked = false;
    m->seg_kernel_ran = false;
    while (done < nsteps || pending) {
        bool launched = false;
        int len = 0;
        tail_checked = false;
        if (done < nsteps) {
            len = std::min(nsteps - done, full_every - m->T % full_every);
            if (!pending) {
                determine_min_max_levels(m);
                if ((rc = upload_tables(m))) return rc;
            }
            if ((rc = launch_steps(m, len, true))) return rc;
            m->T += len;
            done += len;
            launched = true;
            if (pending) *ahead_len = len;
        }
        if (pending) {
            pending = false;
            HIPCHK(hipEventSynchronize(m->ev_check));
            const int J = m->h_changed[2];
            const int ran = J < plen ? J : plen;   // steps the decided segment ran
            st[3] += ran;
            if (J <= plen) st[2]++;                  // the judge found a leaf needing refinement
            const bool at_full = pfull && J >= plen;
            if (at_full) st[0]++;
            if (m->h_changed[0]) {
                HIPCHK(hipStreamSynchronize(m->stream));
                if (launched) {
                    m->T -= len;
                    done -= len;
                    m->segments_skipped++;
                }
                *ahead_len = 0;
                m->T -= plen - ran;
                done -= plen - ran;
                determine_min_max_levels(m);
                m->regrid_changes++;
                if (m->has_net && (rc = enqueue_prediction(m, at_full))) return rc;
                if ((rc = regrid(m, 0.002, 0.001, nullptr, nullptr, m->has_net ? m->d_pred : nullptr, &st[1])))
                    return rc;
                HIPCHK(hipMemsetAsync(m->d_changed, 0, sizeof(int), m->stream));
                continue;
            }
            *ahead_len = 0;
            clear_regrid_flags(m);
            tail_checked = !launched;
        }
        if (launched) {
            plen = len;
            pfull = m->T % full_every == 0;
            if ((rc = enqueue_learned_check(m, len, pfull))) return rc;
            pending = true;
        }
    }
    determine_min_max_levels(m);
    return end_of_steps(m, tail_checked);
}

extern "C" {

int amrlbm_set_regrid_network(amrlbm_mesh *m, const float *weights, double threshold, double gate)
{
    if (!m) return fail(AMRLBM_ERR_ARG, "null mesh");
    if (!(gate >= 0.0)) return fail(AMRLBM_ERR_ARG, "bad prediction gate");
    if (!weights) {
        m->has_net = false;
        return AMRLBM_OK;
    }
    if (!m->d_net_w) {
        HIPCHK(hipMalloc(&m->d_net_w, sizeof(float) * CNN_NPARAM));
        HIPCHK(hipMalloc(&m->d_pred, sizeof(int) * (size_t)m->prm.max_blocks));
        HIPCHK(hipMemsetAsync(m->d_pred, 0, sizeof(int) * (size_t)m->prm.max_blocks, m->stream));
    }
    HIPCHK(hipMemcpyAsync(m->d_net_w, weights, sizeof(float) * CNN_NPARAM, hipMemcpyHostToDevice, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    m->has_net = true;
    m->net_threshold = (float)threshold;
    m->net_gate = gate;
    return AMRLBM_OK;
}

int amrlbm_step_learned(amrlbm_mesh *m, int nsteps, int full_every, long long stats[4])
{
    NEED_READY(m);
    if (nsteps < 0 || full_every < 1 || !stats) return fail(AMRLBM_ERR_ARG, "bad learned-step arguments");
    HIPCHK(hipEventRecord(m->ev0, m->stream));
    int ahead_len = 0;
    const int rc = step_learned(m, nsteps, full_every, stats, &ahead_len);
    if (rc) {
        // as in amrlbm_step: no decision left behind, a skipped segment taken back
        (void)hipStreamSynchronize(m->stream);
        if (ahead_len && *m->h_changed) m->T -= ahead_len;
        (void)hipMemsetAsync(m->d_changed, 0, sizeof(int) * 2, m->stream);
        (void)hipStreamSynchronize(m->stream);
        (void)hipGetLastError();
        return rc;
    }
    return AMRLBM_OK;
}

int amrlbm_step(amrlbm_mesh *m, int nsteps, int regrid_every)
{
    NEED_READY(m);
    if (nsteps < 0 || regrid_every < 1) return fail(AMRLBM_ERR_ARG, "bad step arguments");
    int ahead_len = 0;
    const int rc = step_segments(m, nsteps, regrid_every, &ahead_len);
    if (rc) {
        // an error return leaves no device decision behind: with the stream
        // drained, the skip word every step kernel reads is cleared, so later
        // calls do their work, and a segment launched behind a decision that
        // said "changed" (it skipped its work) is taken back
        (void)hipStreamSynchronize(m->stream);
        if (ahead_len && *m->h_changed) m->T -= ahead_len;
        (void)hipMemsetAsync(m->d_changed, 0, sizeof(int) * 2, m->stream);
        (void)hipStreamSynchronize(m->stream);
        (void)hipGetLastError();
    }
    return rc;
}

int amrlbm_profile_steps(amrlbm_mesh *m, int nsteps, int regrid_every, double stats[12])
{
    NEED_READY(m);
    if (nsteps < 0 || regrid_every < 1 || !stats) return fail(AMRLBM_ERR_ARG, "bad profile arguments");
    for (int k = 0; k < 12; k++) stats[k] = 0.0;
    std::vector<hipEvent_t> ev;
    std::vector<int> kind, cells;
    auto mark = [&](int kd, int nc) -> int {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        HIPCHK(hipEventRecord(e, m->stream));
        ev.push_back(e);
        kind.push_back(kd);
        cells.push_back(nc);
        return AMRLBM_OK;
    };
    auto flush = [&]() -> int {
        HIPCHK(hipStreamSynchronize(m->stream));
        for (size_t k = 1; k < ev.size(); k++) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, ev[k - 1], ev[k]));
            if (cells[k] == 0) continue;
            stats[3 * kind[k]] += ms;
            stats[3 * kind[k] + 1] += 1;
            stats[3 * kind[k] + 2] += cells[k];
        }
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        ev.clear();
        kind.clear();
        cells.clear();
        return AMRLBM_OK;
    };
    int rc = AMRLBM_OK, done = 0;
    while (done < nsteps && !rc) {
        int len = segment_length(m, nsteps - done, regrid_every);
        determine_min_max_levels(m);
        if ((rc = upload_tables(m))) break;
        if ((rc = mark(-1, 0))) break;
        if ((rc = enqueue_segment(m, len, false, mark))) break;
        HIPCHK(hipGetLastError());
        if ((rc = flush())) break;
        m->T += len;
        done += len;
        if (m->T % regrid_every == 0) {
            hipEvent_t a, b;
            HIPCHK(hipEventCreate(&a));
            HIPCHK(hipEventCreate(&b));
            HIPCHK(hipEventRecord(a, m->stream));
            rc = regrid(m, 0.002, 0.001);
            HIPCHK(hipEventRecord(b, m->stream));
            HIPCHK(hipEventSynchronize(b));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, a, b));
            stats[9] += ms;
            stats[10] += 1;
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    if (rc) return rc;
    determine_min_max_levels(m);
    return check_err_flags(m);
}

int amrlbm_time_kernels(amrlbm_mesh *m, int reps, double out[3])
{
    NEED_READY(m);
    if (reps < 1 || !out) return fail(AMRLBM_ERR_ARG, "bad timing arguments");
    int rc = upload_tables(m);
    if (rc) return rc;
    determine_min_max_levels(m);
    const int n = nleaves(m);
    const size_t nb = (size_t)m->nblks;
    // the launches below overwrite populations, C and U,P: keep a copy
    DeviceBuffer buf;
    const size_t e = esz(m), npop = e * POPB * nb, nc = e * CS * nb, npu = e * PUB * nb;
    HIPCHK(hipMalloc(&buf.p, npop + nc + npu));
    char *save = buf.bytes();
    HIPCHK(hipMemcpyAsync(save, m->dm.pop[m->cur], npop, hipMemcpyDeviceToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(save + npop, m->dm.C, nc, hipMemcpyDeviceToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(save + npop + nc, m->dm.PU, npu, hipMemcpyDeviceToDevice, m->stream));
    const int *lst = all_leaves(m);
    for (int kind = 0; kind < 3 && !rc; kind++) {
        auto launch = [&]() {
            StepCtl ctl;
            ctl.out = m->cur ^ 1;   // as in the step: the collision writes the other buffer, the stream writes back
            if (kind == 0)
                launch_prog_c(m->dm, m->cur, lst, n, KIND_C2, OP_MACRO_G_COLLIDE, m->prog[KIND_C2], m->ph, m->stream,
                              nullptr, ctl);
            else if (kind == 1)
                launch_prog_hg(m->dm, m->cur ^ 1, lst, n, m->prog[KIND_HG], true, true, false, m->stream);
            else
                launch_prog_c(m->dm, m->cur, lst, n, KIND_C1, OP_MACRO_G, m->prog[KIND_C1], m->ph, m->stream);
        };
        launch();   // warm
        HIPCHK(hipEventRecord(m->ev0, m->stream));
        for (int r = 0; r < reps; r++) launch();
        HIPCHK(hipEventRecord(m->ev1, m->stream));
        HIPCHK(hipEventSynchronize(m->ev1));
        HIPCHK(hipGetLastError());
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, m->ev0, m->ev1));
        out[kind] = ms / reps;
    }
    HIPCHK(hipMemcpyAsync(m->dm.pop[m->cur], save, npop, hipMemcpyDeviceToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(m->dm.C, save + npop, nc, hipMemcpyDeviceToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(m->dm.PU, save + npop + nc, npu, hipMemcpyDeviceToDevice, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return check_err_flags(m);
}

int amrlbm_set_tuning(amrlbm_mesh *m, const int knobs[3])
{
    if (!m || !knobs) return fail(AMRLBM_ERR_ARG, "null argument");
    if (knobs[0] < 0 || knobs[0] > 3 || knobs[1] < 0 || knobs[1] > 2 || knobs[2] < 0 || knobs[2] > 1)
        return fail(AMRLBM_ERR_ARG, "no such variant");
    if (knobs[0] == m->order && knobs[1] == m->dm.hg_depth && knobs[2] == m->exec) return AMRLBM_OK;
    HIPCHK(hipStreamSynchronize(m->stream));
    m->exec = knobs[2];
    m->seg_lds = 0;   // capacity queried again (another kernel instantiation)
    m->order = knobs[0];
    m->dm.hg_depth = knobs[1];
    drop_graphs(m);   // captured with the previous order and kernels
    m->tables_dirty = true;
    m->topo_version++;
    return m->ready ? upload_tables(m) : AMRLBM_OK;
}

int amrlbm_get_tuning(amrlbm_mesh *m, int knobs[3])
{
    if (!m || !knobs) return fail(AMRLBM_ERR_ARG, "null argument");
    knobs[0] = m->order;
    knobs[1] = m->dm.hg_depth;
    knobs[2] = m->exec;
    return AMRLBM_OK;
}

}  // extern "C"

// Average time of one step of the segment graph of len steps (the step
// kernels as amrlbm_step launches them, no regrid), over reps launches from
// the current state;
// the state is restored afterwards.
static int time_segment(amrlbm_mesh *m, int len, int reps, double *us_per_step)
{
    int rc = upload_tables(m);
    if (rc) return rc;
    determine_min_max_levels(m);
    const size_t nb = (size_t)m->nblks;
    const size_t e = esz(m), npop = e * POPB * nb, nc = e * CS * nb, npu = e * PUB * nb;
    DeviceBuffer buf;
    HIPCHK(hipMalloc(&buf.p, npop + nc + npu));
    char *save = buf.bytes();
    HIPCHK(hipMemcpyAsync(save, m->dm.pop[m->cur], npop, hipMemcpyDeviceToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(save + npop, m->dm.C, nc, hipMemcpyDeviceToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(save + npop + nc, m->dm.PU, npu, hipMemcpyDeviceToDevice, m->stream));
    HIPCHK(hipMemsetAsync(m->d_changed, 0, sizeof(int), m->stream));   // no skip word between regrids
    if ((rc = launch_steps(m, len))) return rc;   // warm
    HIPCHK(hipEventRecord(m->ev0, m->stream));
    for (int r = 0; r < reps && !rc; r++) rc = launch_steps(m, len);
    if (rc) return rc;
    HIPCHK(hipEventRecord(m->ev1, m->stream));
    HIPCHK(hipEventSynchronize(m->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    *us_per_step = 1e3 * ms / ((double)reps * len);
    HIPCHK(hipMemcpyAsync(m->dm.pop[m->cur], save, npop, hipMemcpyDeviceToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(m->dm.C, save + npop, nc, hipMemcpyDeviceToDevice, m->stream));
    HIPCHK(hipMemcpyAsync(m->dm.PU, save + npop + nc, npu, hipMemcpyDeviceToDevice, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return check_err_flags(m);
}

extern "C" {

int amrlbm_autotune(amrlbm_mesh *m, int reps, int regrid_every, double table[96])
{
    NEED_READY(m);
    if (reps < 1 || regrid_every < 1 || !table) return fail(AMRLBM_ERR_ARG, "bad autotune arguments");
    const int len = regrid_every % 2 ? 2 * regrid_every : regrid_every;
    int best[3] = {m->order, m->dm.hg_depth, m->exec};
    double best_us = 1e300;
    int v = 0;
    for (int exec = 0; exec < 2; exec++)
        for (int order = 0; order < 4; order++)
            for (int depth = 0; depth < 3; depth++, v++) {
                const int knobs[3] = {order, depth, exec};
                int rc = amrlbm_set_tuning(m, knobs);
                if (rc) return rc;
                double us = -1.0;   // the segment kernel does not fit: not a variant here
                if ((exec == 0 || use_segment_kernel(m)) && (rc = time_segment(m, len, reps, &us))) return rc;
                table[4 * v] = order;
                table[4 * v + 1] = depth;
                table[4 * v + 2] = exec;
                table[4 * v + 3] = us;
                if (us > 0.0 && us < best_us) {
                    best_us = us;
                    best[0] = order;
                    best[1] = depth;
                    best[2] = exec;
                }
            }
    int rc = amrlbm_set_tuning(m, best);
    if (rc) return rc;
    if (m->exec) return AMRLBM_OK;
    hipGraphExec_t ex;   // the chosen variant's step graph, captured now rather than in the caller's next steps
    return segment_graph(m, len, &ex);
}

// ---- learned regrid predictor --------------------------------------------
namespace {
// device buffers of the predictor entry points (grown on demand, kept until
// the process ends); the entry points run on their own stream
struct CnnCtx {
    hipStream_t s = nullptr;
    float *x = nullptr, *y = nullptr, *logit = nullptr, *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr,
          *loss = nullptr;
    int *idx = nullptr;
    size_t cap = 0;
};
CnnCtx g_cnn;

int cnn_reserve(size_t n)
{
    CnnCtx &c = g_cnn;
    if (!c.s) {
        HIPCHK(hipStreamCreateWithFlags(&c.s, hipStreamNonBlocking));
        HIPCHK(hipMalloc(&c.w, sizeof(float) * CNN_NPARAM));
        HIPCHK(hipMalloc(&c.m, sizeof(float) * CNN_NPARAM));
        HIPCHK(hipMalloc(&c.v, sizeof(float) * CNN_NPARAM));
        HIPCHK(hipMalloc(&c.g, sizeof(float) * CNN_NPARAM));
        HIPCHK(hipMalloc(&c.loss, sizeof(float)));
    }
    if (n > c.cap) {
        (void)hipFree(c.x);
        (void)hipFree(c.y);
        (void)hipFree(c.logit);
        (void)hipFree(c.idx);
        c.x = c.y = c.logit = nullptr;
        c.idx = nullptr;
        c.cap = 0;
        HIPCHK(hipMalloc(&c.x, sizeof(float) * 3 * 18 * 18 * n));
        HIPCHK(hipMalloc(&c.y, sizeof(float) * 2 * n));
        HIPCHK(hipMalloc(&c.logit, sizeof(float) * 2 * n));
        HIPCHK(hipMalloc(&c.idx, sizeof(int) * n));
        c.cap = n;
    }
    return AMRLBM_OK;
}
}  // namespace

int amrlbm_cnn_forward(const float *x, int n, const float *weights, float *logits)
{
    if (!x || !weights || !logits || n < 0) return fail(AMRLBM_ERR_ARG, "bad predictor arguments");
    if (n == 0) return AMRLBM_OK;
    int rc = cnn_reserve(n);
    if (rc) return rc;
    CnnCtx &c = g_cnn;
    HIPCHK(hipMemcpyAsync(c.x, x, sizeof(float) * 3 * 18 * 18 * n, hipMemcpyHostToDevice, c.s));
    HIPCHK(hipMemcpyAsync(c.w, weights, sizeof(float) * CNN_NPARAM, hipMemcpyHostToDevice, c.s));
    launch_cnn(c.x, nullptr, n, c.w, c.logit, nullptr, 0.f, nullptr, nullptr, c.s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(logits, c.logit, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, c.s));
    HIPCHK(hipStreamSynchronize(c.s));
    return AMRLBM_OK;
}

int amrlbm_cnn_gradient(const float *x, const float *y, int n, const float *weights, float pos_weight, float *grad,
                        float *loss)
{
    if (!x || !y || !weights || !grad || !loss || n < 1) return fail(AMRLBM_ERR_ARG, "bad predictor arguments");
    int rc = cnn_reserve(n);
    if (rc) return rc;
    CnnCtx &c = g_cnn;
    HIPCHK(hipMemcpyAsync(c.x, x, sizeof(float) * 3 * 18 * 18 * n, hipMemcpyHostToDevice, c.s));
    HIPCHK(hipMemcpyAsync(c.y, y, sizeof(float) * 2 * n, hipMemcpyHostToDevice, c.s));
    HIPCHK(hipMemcpyAsync(c.w, weights, sizeof(float) * CNN_NPARAM, hipMemcpyHostToDevice, c.s));
    HIPCHK(hipMemsetAsync(c.g, 0, sizeof(float) * CNN_NPARAM, c.s));
    HIPCHK(hipMemsetAsync(c.loss, 0, sizeof(float), c.s));
    launch_cnn(c.x, nullptr, n, c.w, nullptr, c.y, pos_weight, c.g, c.loss, c.s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(grad, c.g, sizeof(float) * CNN_NPARAM, hipMemcpyDeviceToHost, c.s));
    HIPCHK(hipMemcpyAsync(loss, c.loss, sizeof(float), hipMemcpyDeviceToHost, c.s));
    HIPCHK(hipStreamSynchronize(c.s));
    return AMRLBM_OK;
}

int amrlbm_cnn_train(const float *x, const float *y, int n, float *weights, int epochs, int batch, float lr,
                     float pos_weight, unsigned seed, float *epoch_loss)
{
    if (!x || !y || !weights || n < 1 || epochs < 0 || batch < 1 || !(lr > 0.f))
        return fail(AMRLBM_ERR_ARG, "bad training arguments");
    int rc = cnn_reserve(n);
    if (rc) return rc;
    CnnCtx &c = g_cnn;
    HIPCHK(hipMemcpyAsync(c.x, x, sizeof(float) * 3 * 18 * 18 * n, hipMemcpyHostToDevice, c.s));
    HIPCHK(hipMemcpyAsync(c.y, y, sizeof(float) * 2 * n, hipMemcpyHostToDevice, c.s));
    HIPCHK(hipMemcpyAsync(c.w, weights, sizeof(float) * CNN_NPARAM, hipMemcpyHostToDevice, c.s));
    HIPCHK(hipMemsetAsync(c.m, 0, sizeof(float) * CNN_NPARAM, c.s));
    HIPCHK(hipMemsetAsync(c.v, 0, sizeof(float) * CNN_NPARAM, c.s));
    HIPCHK(hipMemsetAsync(c.g, 0, sizeof(float) * CNN_NPARAM, c.s));
    std::vector<int> perm(n);
    for (int k = 0; k < n; k++) perm[k] = k;
    unsigned long long st = seed * 0x9E3779B97F4A7C15ull + 1;   // SplitMix64 shuffle
    auto next = [&st]() {
        unsigned long long z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    int t = 0;
    for (int ep = 0; ep < epochs; ep++) {
        for (int k = n - 1; k > 0; k--) std::swap(perm[k], perm[next() % (unsigned long long)(k + 1)]);
        HIPCHK(hipMemcpyAsync(c.idx, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, c.s));
        HIPCHK(hipMemsetAsync(c.loss, 0, sizeof(float), c.s));
        int nb = 0;
        for (int b0 = 0; b0 < n; b0 += batch, nb++) {
            const int nbat = std::min(batch, n - b0);
            launch_cnn(c.x, c.idx + b0, nbat, c.w, nullptr, c.y, pos_weight, c.g, c.loss, c.s);
            launch_cnn_adam(c.w, c.m, c.v, c.g, lr, ++t, c.s);
        }
        HIPCHK(hipGetLastError());
        float l = 0.f;
        HIPCHK(hipMemcpyAsync(&l, c.loss, sizeof(float), hipMemcpyDeviceToHost, c.s));
        HIPCHK(hipStreamSynchronize(c.s));   // perm is rewritten by the next epoch
        if (epoch_loss) epoch_loss[ep] = l / (float)nb;
    }
    HIPCHK(hipMemcpyAsync(weights, c.w, sizeof(float) * CNN_NPARAM, hipMemcpyDeviceToHost, c.s));
    HIPCHK(hipStreamSynchronize(c.s));
    return AMRLBM_OK;
}

int amrlbm_cnn_features(amrlbm_mesh *m, float *x, int *ids)
{
    NEED_READY(m);
    if (!x || !ids) return fail(AMRLBM_ERR_ARG, "null argument");
    int rc = upload_tables(m);
    if (rc) return rc;
    const int n = nleaves(m);
    float *d = nullptr;
    HIPCHK(hipMalloc(&d, sizeof(float) * 3 * 18 * 18 * std::max(n, 1)));
    launch_cnn_features(m->dm, all_leaves(m), n, d, m->ph.W, m->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(x, d, sizeof(float) * 3 * 18 * 18 * n, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipFree(d));
    for (int k = 0; k < n; k++) ids[k] = m->disp[k];
    return AMRLBM_OK;
}

int amrlbm_output(amrlbm_mesh *m, double out[9])
{
    NEED_READY(m);
    int rc = upload_tables(m);
    if (rc) return rc;
    launch_diag(m->dm, m->d_lists, (int)m->leaf.size(), m->ph, m->d_part, m->d_diag, m->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(m->h_pinned, m->d_diag, sizeof(double) * 9, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    double *d = m->h_pinned;
    for (int k = 0; k < 9; k++) out[k] = d[k];
    out[5] = d[5] / d[4];
    out[6] = d[6] / d[4];
    out[7] = d[7] / d[4];
    if (d[8] != 0.0) return fail(AMRLBM_ERR_NONFINITE, "non-finite phase field in " + std::to_string((long)d[8]) + " cells");
    return AMRLBM_OK;
}

int amrlbm_output_noalloc(amrlbm_mesh *m, double out[9])
{
    NEED_READY(m);
    if (!out) return fail(AMRLBM_ERR_ARG, "null argument");
    if (m->tables_dirty) return fail(AMRLBM_ERR_STATE, "device tables not current (step or regrid first)");
    const int n = (int)m->leaf.size();
    if (n > kScratchLeaves)
        return fail(AMRLBM_ERR_CAPACITY, std::to_string(n) + " leaves: amrlbm_output_noalloc takes at most " +
                                             std::to_string(kScratchLeaves));
    double *part = m->scratch, *tot = m->scratch + kScratchTotals;
    launch_diag(m->dm, m->d_lists, n, m->ph, part, tot, m->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(m->h_pinned, tot, sizeof(double) * 9, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    const double *d = m->h_pinned;
    for (int k = 0; k < 9; k++) out[k] = d[k];
    out[5] = d[5] / d[4];
    out[6] = d[6] / d[4];
    out[7] = d[7] / d[4];
    if (d[8] != 0.0) return fail(AMRLBM_ERR_NONFINITE, "non-finite phase field in " + std::to_string((long)d[8]) + " cells");
    return AMRLBM_OK;
}

int amrlbm_set_mass_correction(amrlbm_mesh *m, int on)
{
    NEED_READY(m);
    if (on != 0 && on != 1) return fail(AMRLBM_ERR_ARG, "the mass correction is 0 (off) or 1 (on)");
    drop_graphs(m);   // the segments' graphs are captured again with or without it
    m->mass_on = 0;
    if (!on) return AMRLBM_OK;
    const int ninst = m->ens ? (int)m->inst.size() : 1;
    if (!m->d_mass_part)
        HIPCHK(hipMalloc(&m->d_mass_part, sizeof(double) * 2 * (size_t)m->prm.max_blocks));
    if (ninst > m->mass_ninst || !m->d_mass) {
        HIPCHK(hipStreamSynchronize(m->stream));
        (void)hipFree(m->d_mass);
        m->d_mass = nullptr;
        HIPCHK(hipMalloc(&m->d_mass, sizeof(double) * 4 * (size_t)ninst));
    }
    m->mass_ninst = ninst;
    int rc = upload_tables(m);   // the device leaf list is current
    if (rc) return rc;
    enqueue_mass_budget(m, MASS_TARGET);
    HIPCHK(hipGetLastError());
    m->mass_on = 1;
    return AMRLBM_OK;
}

int amrlbm_get_mass_budget(amrlbm_mesh *m, int n, double *out)
{
    NEED_READY(m);
    if (!m->mass_on) return fail(AMRLBM_ERR_STATE, "the mass correction is off");
    if (!out || n != m->mass_ninst)
        return fail(AMRLBM_ERR_ARG, "the budget has " + std::to_string(m->mass_ninst) + " instances");
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipMemcpy(out, m->d_mass, sizeof(double) * 4 * (size_t)n, hipMemcpyDeviceToHost));
    return AMRLBM_OK;
}

int amrlbm_synchronize(amrlbm_mesh *m)
{
    if (!m) return fail(AMRLBM_ERR_ARG, "null mesh");
    HIPCHK(hipStreamSynchronize(m->stream));
    return AMRLBM_OK;
}

int amrlbm_last_step_ms(amrlbm_mesh *m, double *ms)
{
    if (!m || !ms) return fail(AMRLBM_ERR_ARG, "null argument");
    *ms = m->last_step_ms;
    return AMRLBM_OK;
}

int amrlbm_regrid_stats(amrlbm_mesh *m, double out[9])
{
    if (!m || !out) return fail(AMRLBM_ERR_ARG, "null argument");
    out[0] = m->regrid_checks;
    out[1] = m->regrid_changes;
    out[2] = m->segments_skipped;
    out[3] = m->host_ms_regrid;
    out[4] = m->host_ms_tables;
    out[5] = m->host_ms_programs;
    out[6] = m->host_ms_graphs;
    out[7] = m->graphs_built;
    out[8] = m->device_flag_regrids;
    return AMRLBM_OK;
}

// ---- per-block kernels on host arrays ------------------------------------
namespace {
struct OneBlock {
    DevMesh dm{};
    BlockDev *blk = nullptr;
    NbDesc *nb = nullptr;
    int *list = nullptr, *flags = nullptr;
    hipStream_t s = nullptr;
    Phys ph{};
    ~OneBlock()
    {
        if (s) (void)hipStreamSynchronize(s);
        (void)hipFree(dm.pop[0]);
        (void)hipFree(dm.pop[1]);
        (void)hipFree(dm.C);
        (void)hipFree(dm.PU);
        (void)hipFree(blk);
        (void)hipFree(nb);
        (void)hipFree(list);
        (void)hipFree(flags);
        if (s) (void)hipStreamDestroy(s);
    }
    int init(const amrlbm_params *p, int dx)
    {
        // dx is L0 / nx / 2^Level, a power of two (the kernels divide by it as a product)
        if (!p || p->nx != NX || p->ny != NY || dx < 1 || (dx & (dx - 1)))
            return fail(AMRLBM_ERR_ARG, "bad per-block arguments (16 x 16 blocks, dx a power of two)");
        ph.M0 = p->M0; ph.wc = 1.0 / (0.5 + p->M0 * 3.0); ph.W = p->W; ph.sigma = p->sigma; ph.beta = p->beta;
        ph.kappa = p->kappa; ph.taul = p->taul; ph.tauh = p->tauh; ph.rhol = p->rhol; ph.rhoh = p->rhoh;
        ph.gy = p->gy;
        ph.drho3 = (p->rhoh - p->rhol) / 3.0;
        ph.drho6 = (p->rhoh - p->rhol) / 6.0;
        ph.inv_tauh = 1.0 / p->tauh;
        ph.inv_taul = 1.0 / p->taul;
        ph.inv_W = 1.0 / p->W;
        HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        HIPCHK(hipMalloc(&dm.pop[0], sizeof(double) * POPB));
        HIPCHK(hipMalloc(&dm.pop[1], sizeof(double) * POPB));
        HIPCHK(hipMalloc(&dm.C, sizeof(double) * CS));
        HIPCHK(hipMalloc(&dm.PU, sizeof(double) * PUB));
        HIPCHK(hipMalloc(&blk, sizeof(BlockDev)));
        HIPCHK(hipMalloc(&nb, sizeof(NbDesc) * 8));
        HIPCHK(hipMalloc(&list, sizeof(int)));
        HIPCHK(hipMalloc(&flags, sizeof(int) * 2));   // [0] error flags, [1] skip word (zero)
        HIPCHK(hipMemsetAsync(dm.pop[0], 0, sizeof(double) * POPB, s));
        HIPCHK(hipMemsetAsync(dm.pop[1], 0, sizeof(double) * POPB, s));
        HIPCHK(hipMemsetAsync(dm.C, 0, sizeof(double) * CS, s));
        HIPCHK(hipMemsetAsync(dm.PU, 0, sizeof(double) * PUB, s));
        HIPCHK(hipMemsetAsync(list, 0, sizeof(int), s));
        HIPCHK(hipMemsetAsync(flags, 0, sizeof(int) * 2, s));
        BlockDev b{0, dx, 0, 0, 1, -1, 0, 0};
        NbDesc nbh[8];
        for (int k = 0; k < 8; k++) nbh[k] = NbDesc{-(k + 1), -1, 0, -1};
        HIPCHK(hipMemcpyAsync(blk, &b, sizeof(b), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(nb, nbh, sizeof(nbh), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        dm.blk = blk;
        dm.nbd = nb;
        dm.err_flags = flags;
        dm.skip = flags + 1;
        return AMRLBM_OK;
    }
    int put_pop(int set, const double *ref)
    {
        std::vector<double> buf(9 * GS);
        ref_to_dev_pop(ref, buf.data());
        HIPCHK(hipMemcpyAsync(dm.pop[0] + set * 9 * GS, buf.data(), sizeof(double) * 9 * GS, hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        return AMRLBM_OK;
    }
    int get_pop(int b, int set, double *ref)
    {
        std::vector<double> buf(9 * GS);
        HIPCHK(hipMemcpyAsync(buf.data(), dm.pop[b] + set * 9 * GS, sizeof(double) * 9 * GS, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        dev_to_ref_pop(buf.data(), ref);
        return AMRLBM_OK;
    }
};
}  // namespace

int amrlbm_collision_block(const amrlbm_params *p, int dx, const double *C, const double *P, const double *Ux,
                           const double *Uy, double *h, double *g)
{
    if (!C || !P || !Ux || !Uy || !h || !g) return fail(AMRLBM_ERR_ARG, "null argument");
    OneBlock o;
    int rc = o.init(p, dx);
    if (rc) return rc;
    if ((rc = o.put_pop(0, h)) || (rc = o.put_pop(1, g))) return rc;
    HIPCHK(hipMemcpyAsync(o.dm.C, C, sizeof(double) * CS, hipMemcpyHostToDevice, o.s));
    HIPCHK(hipMemcpyAsync(o.dm.PU, P, sizeof(double) * PS, hipMemcpyHostToDevice, o.s));
    HIPCHK(hipMemcpyAsync(o.dm.PU + PS, Ux, sizeof(double) * PS, hipMemcpyHostToDevice, o.s));
    HIPCHK(hipMemcpyAsync(o.dm.PU + 2 * PS, Uy, sizeof(double) * PS, hipMemcpyHostToDevice, o.s));
    launch_prog_c(o.dm, 0, o.list, 1, KIND_C2, OP_COLLIDE, kNoProg, o.ph, o.s);
    HIPCHK(hipGetLastError());
    if ((rc = o.get_pop(0, 0, h)) || (rc = o.get_pop(0, 1, g))) return rc;
    return AMRLBM_OK;
}

int amrlbm_macro_g_block(const amrlbm_params *p, int dx, const double *g, const double *C, double *P, double *Ux,
                         double *Uy)
{
    if (!g || !C || !P || !Ux || !Uy) return fail(AMRLBM_ERR_ARG, "null argument");
    OneBlock o;
    int rc = o.init(p, dx);
    if (rc) return rc;
    if ((rc = o.put_pop(1, g))) return rc;
    HIPCHK(hipMemcpyAsync(o.dm.C, C, sizeof(double) * CS, hipMemcpyHostToDevice, o.s));
    launch_prog_c(o.dm, 0, o.list, 1, KIND_C1, OP_MACRO_G, kNoProg, o.ph, o.s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(P, o.dm.PU, sizeof(double) * PS, hipMemcpyDeviceToHost, o.s));
    HIPCHK(hipMemcpyAsync(Ux, o.dm.PU + PS, sizeof(double) * PS, hipMemcpyDeviceToHost, o.s));
    HIPCHK(hipMemcpyAsync(Uy, o.dm.PU + 2 * PS, sizeof(double) * PS, hipMemcpyDeviceToHost, o.s));
    HIPCHK(hipStreamSynchronize(o.s));
    return AMRLBM_OK;
}

int amrlbm_stream_block(const amrlbm_params *p, int dx, double *f)
{
    if (!f) return fail(AMRLBM_ERR_ARG, "null argument");
    OneBlock o;
    int rc = o.init(p, dx);
    if (rc) return rc;
    if ((rc = o.put_pop(0, f))) return rc;
    launch_prog_hg(o.dm, 0, o.list, 1, kNoProg, true, false, true, o.s);
    HIPCHK(hipGetLastError());
    return o.get_pop(1, 0, f);
}

}  // extern "C"
