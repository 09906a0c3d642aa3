// mgp_api.cpp -- host orchestration and the C ABI of libmgp.so.
//
// Mirrors the Fortran interface of src/flow/pressuresolver_mod.F90:
// mgp_init ~ init_pressuresolver (+ ib%giteig), mgp_mgpoisl ~ mgpoisl,
// mgp_finish ~ finish_pressuresolver.  All work is issued on one HIP
// stream per handle; the outer-iteration loop is captured into a HIP graph
// and its exit test runs on the device, so the host only polls a flag every
// `check_every` outer iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <memory>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/mgp.h"
#include "mgp_launch.h"

using namespace mgp;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

struct HipError {
    std::string msg;
};

void hip_check(hipError_t e, const char *what, int line) {
    if (e != hipSuccess) {
        char buf[512];
        snprintf(buf, sizeof buf, "%s failed (line %d): %s", what, line,
                 hipGetErrorString(e));
        throw HipError{buf};
    }
}
#define HC(x) hip_check((x), #x, __LINE__)

// device buffer owned by a handle
struct DBuf {
    void *p = nullptr;
    size_t n = 0;
    ~DBuf() { if (p) (void)hipFree(p); }
    void alloc(size_t bytes) {
        if (p) { (void)hipFree(p); p = nullptr; }
        n = bytes;
        if (bytes) HC(hipMalloc(&p, bytes));
    }
    template <typename T> T *as() const { return (T *)p; }
};

// Faces and directions (src/core/connect/connect_core_mod.F90:37-165): the
// face numbering is part of the data format.
const int FACELIST[26][4] = {
    {1, 1, 0, 0}, {1, 2, 0, 0}, {1, 3, 0, 0}, {1, 4, 0, 0}, {1, 5, 0, 0},
    {1, 6, 0, 0}, {2, 1, 3, 0}, {2, 1, 4, 0}, {2, 1, 5, 0}, {2, 1, 6, 0},
    {2, 2, 3, 0}, {2, 2, 4, 0}, {2, 2, 5, 0}, {2, 2, 6, 0}, {2, 3, 5, 0},
    {2, 3, 6, 0}, {2, 4, 5, 0}, {2, 4, 6, 0}, {3, 1, 3, 5}, {3, 1, 3, 6},
    {3, 1, 4, 5}, {3, 1, 4, 6}, {3, 2, 3, 5}, {3, 2, 3, 6}, {3, 2, 4, 5},
    {3, 2, 4, 6}};
const int FACENBR[26] = {2,  1,  4,  3,  6,  5,  12, 11, 14, 13, 8,  7,  10,
                         9,  18, 17, 16, 15, 26, 25, 24, 23, 22, 21, 20, 19};
// neighbour search order per receiving direction and the face it is sent
// from (rescue_dir / rescue_nbr, connect_core_mod.F90:76-151)
const int RESCUE_DIR[26][7] = {
    {1}, {2}, {3}, {4}, {5}, {6}, {7, 1, 3}, {8, 4, 1}, {9, 1, 5},
    {10, 6, 1}, {11, 2, 3}, {12, 2, 4}, {13, 2, 5}, {14, 2, 6}, {15, 3, 5},
    {16, 6, 3}, {17, 4, 5}, {18, 4, 6}, {19, 7, 9, 15, 1, 3, 5},
    {20, 7, 10, 16, 1, 3, 6}, {21, 8, 9, 17, 1, 4, 5},
    {22, 8, 10, 18, 1, 4, 6}, {23, 11, 13, 15, 2, 3, 5},
    {24, 11, 14, 16, 2, 3, 6}, {25, 12, 13, 17, 2, 4, 5},
    {26, 12, 14, 18, 2, 4, 6}};
const int RESCUE_NBR[26][7] = {
    {2}, {1}, {4}, {3}, {6}, {5}, {12, 11, 8}, {11, 7, 12}, {14, 13, 10},
    {13, 9, 14}, {8, 7, 12}, {7, 8, 11}, {10, 9, 14}, {9, 10, 13},
    {18, 17, 16}, {17, 15, 18}, {16, 15, 18}, {15, 16, 17},
    {26, 25, 24, 22, 23, 21, 20}, {25, 26, 23, 21, 24, 22, 19},
    {24, 23, 26, 20, 25, 19, 22}, {23, 24, 25, 19, 26, 20, 21},
    {22, 21, 20, 26, 19, 25, 24}, {21, 22, 19, 25, 20, 26, 23},
    {20, 19, 22, 24, 21, 23, 26}, {19, 20, 21, 23, 22, 24, 25}};

bool is_con(int c) { return c == 7 || c == 19; }
constexpr int PAR_BC = 8;
// flow boundary codes bound_flow acts on (src/flow/bound_flow_mod.F90)
constexpr int FIX_BC = 2, OP1_BC = 3, NOS_BC = 5, SLI_BC = 6;
bool is_flow_bc(int c) {
    return c == FIX_BC || c == OP1_BC || c == NOS_BC || c == SLI_BC || c == PAR_BC;
}
bool is_wall_lo(int c) { return c == 2 || c == 5 || c == 6 || c == 19; }
bool is_wall_hi(int c) { return c == 2 || c == 5 || c == 6; }

struct Conn {
    int recv, send, frecv, fsend;  // 1-based grid ids, face numbers
};

struct Box {
    int is, ie, js, je, ks, ke;
};

}  // namespace

// Scratch of the allocation-free entry point mgp_divcal_device_noalloc: the
// divergence in WO order before it is scattered into the caller's field.
// One array per GPU, held until the process ends.  Its host-side name is a
// stand-in, not a device address: host code passes the address that
// hipGetSymbolAddress returns for the current device (scratch_device_ptr).
constexpr size_t kScratchCells = size_t(1) << 23;   // 8 Mi FP64 values, 64 MiB
static __device__ __align__(256) double scratch[kScratchCells];

namespace {

// Device address of `scratch` on the current device, looked up once per
// device: each GPU holds its own copy at its own address.
void *scratch_device_ptr() {
    static std::mutex mu;
    static std::map<int, void *> ptrs;
    int dev = 0;
    HC(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = ptrs.find(dev);
    if (it != ptrs.end()) return it->second;
    void *ptr = nullptr;
    HC(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw HipError{"hipGetSymbolAddress returned no address for the scratch array"};
    ptrs[dev] = ptr;
    return ptr;
}

}  // namespace

// ---------------------------------------------------------------------
struct mgp_handle {
    // configuration and topology (host)
    mgp_config cfg{};
    int ngrid = 0, minlevel = 0, maxlevel = 0;
    std::vector<int> kk, jj, ii, level, nbr, bc;
    std::vector<int64_t> ip3, ip1x, ip1y, ip1z;  // 0-based
    std::vector<int> iparent, position;         // 1-based parent ids, positions
    int64_t n3d = 0, n1x = 0, n1y = 0, n1z = 0, ncell = 0, nface = 0, nshell = 0;
    std::vector<Shape> shapes;
    std::vector<Block> blocks;
    std::vector<Conn> conns;
    std::map<int, std::vector<int>> level_blocks;  // level -> block indices
    bool masked = false;
    int max_ni = 0, max_nj = 0, max_nk = 0, max_plane = 0;

    hipStream_t own_stream = nullptr, stream = nullptr;

    // device topology
    DBuf d_blocks, d_shapes, d_cells, d_pstart, d_slot, d_blist_all;
    DBuf d_items, d_ents, d_zero, d_cellx, d_fents;   // plans of the plane kernels
    std::map<int, DBuf> d_blist_lvl;
    // 1-D fields (FP64 and working precision)
    DBuf d_dx, d_dy, d_dz, d_rdx, d_rdy, d_rdz, d_rddx, d_rddy, d_rddz;
    DBuf d_ddx, d_ddy, d_ddz;
    DBuf d_rpd[3];       // 1 / (rho dx), 1 / (rho dy), 1 / (rho dz) (tstle4)
    DBuf d_raw[6], d_c64[6], d_cT[6], d_apd[3];
    // per-cell WO arrays
    DBuf d_ap64, d_apT, d_f[7], d_hilf, d_rhs64, d_rhsT, d_res, d_dp64;
    // shared factor storage: classes of blocks with equal factors, their
    // representatives (device list) and the factor cells stored
    int nfclass = 0;
    int64_t nfcell = 0;
    std::vector<int> frep;
    DBuf d_blist_rep;
    DBuf d_aps64, d_apsT;      // GSAP kept with the shared factors (FP64, FP32)
    bool fshared() const { return d_aps64.p != nullptr; }
    DBuf d_bpw, d_bpf, d_bpsp;
    DBuf d_sdiv;               // ghostcell divergence source (reference layout)
    DBuf d_hface, d_rhi[2], d_rlo;
    // SOR: SOR_RAP and the second HILF face buffer (frozen ghosts)
    DBuf d_rapT, d_hface2;
    // levels: PAR ghosts of HILF, parent values per PAR face group (HILF in
    // the working precision, DP in FP64)
    DBuf d_hpar, d_pbuf, d_pbuf64;
    // first ghost layer ("shell") of refined grids: HILF as ctof injects it
    // (read by the next level's ctof), and its sum over outer iterations,
    // which DP keeps in the ghost cells no exchange or bound_pressure sets
    DBuf d_hsh, d_dsh;
    std::set<int> level_par;   // levels with PAR faces
    DBuf d_parlist;            // blocks with PAR faces (tstle4_par)
    int npar = 0;
    // par_ftoc_norm transfers, ordered by the fine grid's level, and the
    // range of each level in that table
    DBuf d_pftask;
    std::map<int, std::pair<int, int>> pf_range;
    // closing sync of mgpoisl: velocity boundary buffers of the FIX, OP1
    // and PAR faces (offsets per grid and face, -1 none), OP1 pressure per
    // grid and face, and the levels with a face bound_flow acts on
    std::vector<int64_t> bbo;
    int64_t nbb = 0;
    DBuf d_bbo, d_ubnd, d_vbnd, d_wbnd, d_pinf;
    std::set<int> level_flow;
    // adjoint: transposed factors and 1-D coefficients (built on first use)
    DBuf d_ft[6], d_c64t[6], d_cTt[6];
    std::vector<double> fold_host[6];
    bool adj_ready = false;
    // reference-layout work arrays
    DBuf d_u, d_v, d_w, d_p, d_dpsp;
    // this device's copy of the static scratch array (set at init)
    double *scratch = nullptr;
    // scalars and history
    DBuf d_scal, d_hist, d_never;
    // exchange tasks: key (level, layers, corners, forward)
    struct TaskSet {
        DBuf blocks, ents;             // PullBlock records of all stages, entry tables
        std::vector<int> stage_start;  // per stage offsets into blocks (+ end)
        bool pairs = false;            // entries move two values along k
        bool valid = true;             // false: a kept copy touches a pushed region
    };
    std::map<std::tuple<int, int, int, int, int>, std::unique_ptr<TaskSet>> tasks;
    // mgpoisl's velocity and pressure correction along z-chains, which
    // writes the k-face ghosts of interior rows itself (single-level grids):
    // 0 not decided, 1 on, -1 off; the chains (CorrChains) and per grid the
    // faces whose interior-row ghosts it writes (bit 0 k-low, bit 1 k-high)
    int kpush_state = 0;
    DBuf d_cchain, d_cchains;
    int ncchain = 0, ccnx = 0;
    std::vector<int> kpush_faces;
    // captured outer-iteration graphs per level
    hipGraphExec_t graph = nullptr;
    int graph_key[4] = {-1, -1, -1, -1};
    bool have_solution = false;
    int plan_shape = 1;          // pipeline shape of the plane kernels
    int plan_shape_bwd = 1;      // ... of the backward sweep
    double shape_ms[kNumShapes] = {};
    static_assert(kNumShapes == MGP_NUM_PLAN_SHAPES, "shape count of the C ABI");

    ~mgp_handle() {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (own_stream) (void)hipStreamDestroy(own_stream);
    }

    bool fp32() const { return cfg.precision == 1; }
    const double *sdiv_sp() const {
        return cfg.ibtype == 1 && d_sdiv.p ? d_sdiv.as<double>() : nullptr;
    }
    size_t tsz() const { return fp32() ? 4 : 8; }

    Topo topo(int lvl) const {
        Topo t;
        t.blocks = d_blocks.as<Block>();
        t.shapes = d_shapes.as<Shape>();
        t.cells = d_cells.as<uint32_t>();
        t.pstart = d_pstart.as<int>();
        t.slot_of = d_slot.as<int>();
        t.max_ni = max_ni;
        t.max_nj = max_nj;
        t.max_nk = max_nk;
        t.max_plane = max_plane;
        t.maxlevel = maxlevel;
        t.items = d_items.as<uint2>();
        t.ents = d_ents.as<uint32_t>();
        t.cellx = d_cellx.as<uint32_t>();
        t.fents = d_fents.as<uint32_t>();
        t.zero = d_zero.p;
        if (lvl < 0) {
            t.blist = d_blist_all.as<int>();
            t.nlist = ngrid;
        } else {
            auto it = d_blist_lvl.find(lvl);
            t.blist = it->second.as<int>();
            t.nlist = (int)level_blocks.at(lvl).size();
        }
        return t;
    }

    // the representatives of the factor classes (one block per class)
    Topo topo_reps() const {
        if (d_blist_rep.n != (size_t)nfclass * sizeof(int))
            throw HipError{"factor class list not built"};
        Topo t = topo(-1);
        t.blist = d_blist_rep.as<int>();
        t.nlist = nfclass;
        return t;
    }

    Mask mask() const {
        Mask m;
        m.bp = masked ? d_bpw.as<uint8_t>() : nullptr;
        m.bpf = masked ? d_bpf.as<uint8_t>() : nullptr;
        return m;
    }

    double *scal() const { return d_scal.as<double>(); }
    Outer outer() const {
        Outer o;
        double *s = scal();
        o.rhs64 = d_rhs64.as<double>();
        o.dp64 = d_dp64.as<double>();
        o.ap64 = d_ap64.as<double>();
        o.aw = d_c64[0].as<double>();
        o.ae = d_c64[1].as<double>();
        o.as = d_c64[2].as<double>();
        o.an = d_c64[3].as<double>();
        o.ab = d_c64[4].as<double>();
        o.at = d_c64[5].as<double>();
        o.gmax = (unsigned long long *)(s + 0);
        o.maxrhs0 = s + 1;
        o.maxrhs = s + 2;
        int *iv = (int *)(s + 3);
        o.ctr = iv + 0;
        o.done = iv + 1;
        o.converged = iv + 2;
        o.nonfinite = iv + 3;
        return o;
    }

    template <typename T>
    Fields<T> fields() const {
        Fields<T> F;
        F.hilf = d_hilf.as<T>();
        F.rhs = fp32() ? d_rhsT.as<T>() : (T *)d_rhs64.as<T>();
        F.res = d_res.as<T>();
        F.ap = fp32() ? d_apT.as<T>() : (T *)d_ap64.as<T>();
        for (int q = 0; q < 7; ++q) F.f[q] = d_f[q].as<T>();
        const T **c[6] = {&F.aw, &F.ae, &F.as, &F.an, &F.ab, &F.at};
        for (int q = 0; q < 6; ++q)
            *c[q] = fp32() ? d_cT[q].as<T>() : d_c64[q].as<T>();
        F.rap = d_rapT.as<T>();
        F.apx = d_apd[0].as<double>();
        F.apy = d_apd[1].as<double>();
        F.apz = d_apd[2].as<double>();
        F.aps = !fshared() ? nullptr : fp32() ? d_apsT.as<T>() : d_aps64.as<T>();
        F.hface = d_hface.as<T>();
        F.hpar = d_hpar.as<T>();
        F.rhi[0] = d_rhi[0].as<T>();
        F.rhi[1] = d_rhi[1].as<T>();
        F.rlo = d_rlo.as<T>();
        return F;
    }

    // operands of A^T: transposed factors in the slots the sweeps read
    // (LWt LSt LBt LPR UEt UNt UTt) and the transposed 1-D coefficients
    template <typename T>
    Fields<T> fields_t() const {
        Fields<T> F = fields<T>();
        F.f[0] = d_ft[0].as<T>();
        F.f[1] = d_ft[1].as<T>();
        F.f[2] = d_ft[2].as<T>();
        F.f[4] = d_ft[3].as<T>();
        F.f[5] = d_ft[4].as<T>();
        F.f[6] = d_ft[5].as<T>();
        const T **c[6] = {&F.aw, &F.ae, &F.as, &F.an, &F.ab, &F.at};
        for (int q = 0; q < 6; ++q)
            *c[q] = fp32() ? d_cTt[q].as<T>() : d_c64t[q].as<T>();
        return F;
    }

    Outer outer_t() const {
        Outer o = outer();
        o.aw = d_c64t[0].as<double>();
        o.ae = d_c64t[1].as<double>();
        o.as = d_c64t[2].as<double>();
        o.an = d_c64t[3].as<double>();
        o.ab = d_c64t[4].as<double>();
        o.at = d_c64t[5].as<double>();
        return o;
    }
};

namespace {

// ------------------------------------------------------------------
// Connection tables: src/core/connect/connect_core_mod.F90:167-433.
void get_nbrs(const mgp_handle &h, int g, int iface, int &ng, int &nf) {
    ng = nf = 0;
    const int *nb = &h.nbr[26 * (g - 1)];
    for (int r = 0; r < 7; ++r) {
        const int dir = RESCUE_DIR[iface - 1][r];
        if (dir == 0) break;
        if (nb[dir - 1] > 0) {
            ng = nb[dir - 1];
            nf = RESCUE_NBR[iface - 1][r];
            if (nf > 18) {
                const int *bcn = &h.bc[6 * (ng - 1)];
                bool any = false;
                for (int a = 1; a <= 3; ++a)
                    any = any || is_con(bcn[FACELIST[nf - 1][a] - 1]);
                if (!any) {
                    ng = nf = 0;
                    continue;
                }
            }
            break;
        }
    }
}

// Transfers of par_ftoc_norm (the connection table of init_par_ftoc,
// src/ib/par_ftoc_mod.F90:513-745): every PAR face of a fine grid sends to
// its parent, at the parent's own boundary face when the fine face lies on
// it (the fine grid starts at the parent's first cell, pos == 3, for a low
// face; ends at its last, for a high face: pos > 3, or pos == 3 with the
// fine grid spanning the whole parent) and otherwise inside (the coarse
// face before the first / after the last covered cell); a face on the
// parent's boundary also goes to the parent's neighbour across it (NBRGRID
// of the parent), at that grid's opposite boundary face.
template <typename T>
void upload(DBuf &d, const std::vector<T> &v);

void build_par_ftoc(mgp_handle &h) {
    std::vector<ParFtocTask> tasks;
    for (int L = h.minlevel; L <= h.maxlevel; ++L) {
        const int first = (int)tasks.size();
        for (int g = 0; g < h.ngrid; ++g) {
            const Block &B = h.blocks[g];
            if (B.level != L || !B.par) continue;
            const int c = B.parent;
            const int nf[3] = {h.ii[g], h.jj[g], h.kk[g]};
            const int nc[3] = {h.ii[c], h.jj[c], h.kk[c]};
            for (int f = 0; f < 6; ++f) {
                if (!((B.par >> f) & 1)) continue;
                const int ax = f >> 1, pos = B.pos[ax];
                const bool low = (f & 1) == 0, one_connect = !(nf[ax] > nc[ax]);
                bool inside = false;
                if (pos == 3 && one_connect) inside = !low;
                else if (pos > 3) inside = low;
                ParFtocTask t;
                t.fine = g;
                t.face = f;
                t.dest = c;
                if (inside) t.dnorm = low ? pos - 1 : pos + (nf[ax] - 4) / 2 - 1;
                else t.dnorm = low ? 2 : nc[ax] - 2;
                tasks.push_back(t);
                if (inside) continue;
                const int nb = h.nbr[26 * c + f];   // 1-based, 0 none
                if (nb <= 0) continue;
                const int nn[3] = {h.ii[nb - 1], h.jj[nb - 1], h.kk[nb - 1]};
                t.dest = nb - 1;
                t.dnorm = low ? nn[ax] - 2 : 2;
                tasks.push_back(t);
            }
        }
        if ((int)tasks.size() > first) h.pf_range[L] = {first, (int)tasks.size() - first};
    }
    if (!tasks.empty()) upload(h.d_pftask, tasks);
}

void build_conns(mgp_handle &h) {
    h.conns.clear();
    for (int g = 1; g <= h.ngrid; ++g) {
        const int *bc = &h.bc[6 * (g - 1)];
        for (int f = 1; f <= 26; ++f) {
            bool want = false;
            for (int a = 1; a <= FACELIST[f - 1][0]; ++a)
                want = want || is_con(bc[FACELIST[f - 1][a] - 1]);
            if (!want) continue;
            int ng, nf;
            get_nbrs(h, g, f, ng, nf);
            if (ng == 0) continue;
            h.conns.push_back({g, ng, f, nf});
        }
    }
}

// start_and_stop(_face) / corr_start_stop, connect_core_mod.F90:454-602
void ss_face(const mgp_handle &h, int g, int f, Box &b, int np, bool ng) {
    const int kk = h.kk[g - 1], jj = h.jj[g - 1], ii = h.ii[g - 1];
    switch (f) {
    case 1: b.is = 3 - np; b.ie = 2; break;
    case 2: b.is = ii - 1; b.ie = ii - (2 - np); break;
    case 3: b.js = 3 - np; b.je = 2; break;
    case 4: b.js = jj - 1; b.je = jj - (2 - np); break;
    case 5: b.ks = 3 - np; b.ke = 2; break;
    case 6: b.ks = kk - 1; b.ke = kk - (2 - np); break;
    }
    if (ng) {
        switch (f) {
        case 1: b.is += np; b.ie += np; break;
        case 2: b.is -= np; b.ie -= np; break;
        case 3: b.js += np; b.je += np; break;
        case 4: b.js -= np; b.je -= np; break;
        case 5: b.ks += np; b.ke += np; break;
        case 6: b.ks -= np; b.ke -= np; break;
        }
    }
}

Box ss(const mgp_handle &h, int g, int f, int np, bool ng) {
    Box b{3, h.ii[g - 1] - 2, 3, h.jj[g - 1] - 2, 3, h.kk[g - 1] - 2};
    for (int a = 1; a <= FACELIST[f - 1][0]; ++a)
        ss_face(h, g, FACELIST[f - 1][a], b, np, ng);
    return b;
}

bool overlap(const Box &a, const Box &b) {
    return a.is <= b.ie && b.is <= a.ie && a.js <= b.je && b.js <= a.je &&
           a.ks <= b.ke && b.ks <= a.ke;
}

template <typename T>
void upload(DBuf &d, const std::vector<T> &v);

// The grids of a set (1-based ids, ascending) in nested neighbour chains:
// along +x (face 2) outermost, then +y (face 4), then +z (face 6)
// innermost, so that consecutive entries are z-neighbours and
// neighbouring columns y-neighbours.  Every grid starts at most one walk
// per direction: O(ngrid).
std::vector<int> chain_order(const mgp_handle &h, const std::vector<int> &ids) {
    std::vector<char> in(h.ngrid + 1, 0), zdone(h.ngrid + 1, 0), ydone(h.ngrid + 1, 0),
        xdone(h.ngrid + 1, 0);
    for (int g : ids) in[g] = 1;
    std::vector<int> order;
    auto next = [&](int g, int face) { return h.nbr[26 * (g - 1) + face - 1]; };
    auto zwalk = [&](int b) {
        for (int g = b; g > 0 && !zdone[g]; g = next(g, 6)) {
            zdone[g] = 1;
            if (in[g]) order.push_back(g);
        }
    };
    auto ywalk = [&](int a) {
        for (int g = a; g > 0 && !ydone[g]; g = next(g, 4)) {
            ydone[g] = 1;
            zwalk(g);
        }
    };
    for (int g0 : ids)
        for (int g = g0; g > 0 && !xdone[g]; g = next(g, 2)) {
            xdone[g] = 1;
            ywalk(g);
        }
    if (order.size() != ids.size()) throw HipError{"block chain order incomplete"};
    return order;
}

// Copy tasks of one exchange, grouped into stages that keep the
// reference's sequential order wherever two copies touch the same cells
// (only possible when a source box reaches into a ghost layer).
// kpush: the k-face copies into the interior-row ghosts that the pushing
// velocity and pressure correction fills (h.kpush_faces) are left out.
mgp_handle::TaskSet &task_set(mgp_handle &h, int level, int layers,
                              int corners, int forward, bool kpush = false) {
    auto key = std::make_tuple(level, layers, corners, forward, (int)kpush);
    auto it = h.tasks.find(key);
    if (it != h.tasks.end()) return *it->second;
    const int minl = level >= 0 ? level : h.minlevel;
    const int maxl = level >= 0 ? level : h.maxlevel;
    struct T2 {
        CopyTask t;
        int sg, dg;
        Box sb, db;
    };
    std::vector<T2> list;
    std::vector<std::vector<Box>> pushed(kpush ? h.ngrid + 1 : 0);   // per grid, left out
    for (const Conn &c : h.conns) {
        const int fr = c.frecv;
        const int lvl = h.level[c.recv - 1];
        if (lvl > maxl || lvl < minl) continue;          // decide()
        if (!corners && fr > 6) continue;
        if (forward == 1 && (fr == 2 || fr == 4 || fr == 6)) continue;
        if (forward == -1 && (fr == 1 || fr == 3 || fr == 5)) continue;
        if (kpush && (fr == 5 || fr == 6) &&
            (h.kpush_faces[c.recv - 1] >> (fr - 5) & 1)) {
            pushed[c.recv].push_back(ss(h, c.recv, fr, layers, false));
            continue;
        }
        Box sb = ss(h, c.send, FACENBR[fr - 1], layers, true);
        if (fr > 6)
            for (int a = 1; a <= FACELIST[c.fsend - 1][0]; ++a)
                for (int q = 1; q <= FACELIST[fr - 1][0]; ++q)
                    if (FACELIST[c.fsend - 1][a] == FACELIST[fr - 1][q])
                        ss_face(h, c.send, FACELIST[c.fsend - 1][a], sb,
                                layers, false);
        Box db = ss(h, c.recv, fr, layers, false);
        const int ni = db.ie - db.is + 1, nj = db.je - db.js + 1,
                  nk = db.ke - db.ks + 1;
        if ((sb.ie - sb.is + 1) * (sb.je - sb.js + 1) * (sb.ke - sb.ks + 1) !=
            ni * nj * nk)
            throw HipError{"inconsistent connection box sizes"};
        const int64_t kks = h.kk[c.send - 1], jjs = h.jj[c.send - 1];
        const int64_t kkd = h.kk[c.recv - 1], jjd = h.jj[c.recv - 1];
        CopyTask t;
        t.src = h.ip3[c.send - 1] + (sb.ks - 1) + (sb.js - 1) * kks +
                (sb.is - 1) * kks * jjs;
        t.dst = h.ip3[c.recv - 1] + (db.ks - 1) + (db.js - 1) * kkd +
                (db.is - 1) * kkd * jjd;
        t.ni = ni; t.nj = nj; t.nk = nk;
        t.s_sj = (int)kks; t.s_si = (int)(kks * jjs);
        t.d_sj = (int)kkd; t.d_si = (int)(kkd * jjd);
        list.push_back({t, c.send, c.recv, sb, db});
    }
    // a kept copy that reads or writes a pushed ghost region would see it
    // filled out of the reference's order
    bool valid = true;
    if (kpush)
        for (const T2 &q : list) {
            for (const Box &b : pushed[q.sg]) valid = valid && !overlap(q.sb, b);
            for (const Box &b : pushed[q.dg]) valid = valid && !overlap(q.db, b);
        }
    // A source box that reaches into a destination box of its own grid (a
    // two-layer exchange of blocks one cell thick along an axis, whose
    // source boxes take a ghost layer, also a copy's own destination) makes
    // the result depend on the element order inside a copy.
    std::vector<std::vector<int>> wr(h.ngrid + 1), rdl(h.ngrid + 1);
    bool hazard = false;
    for (size_t q = 0; q < list.size(); ++q) wr[list[q].dg].push_back((int)q);
    for (size_t q = 0; q < list.size() && !hazard; ++q)
        for (int u : wr[list[q].sg])
            if (overlap(list[u].db, list[q].sb)) { hazard = true; break; }
    if (hazard) {
        // Then the copies run one element at a time in the reference's
        // order (connections in turn; i, j, k with k fastest) on the host:
        // each written position ends with the value of one original
        // position, and an entry that reads a position another entry writes
        // runs in an earlier stage (longest chain).
        struct Orig { int64_t pos; int grid; };
        std::unordered_map<int64_t, Orig> cur;
        std::vector<std::pair<int64_t, int>> written;   // position, receiving grid
        for (const T2 &q : list) {
            const CopyTask &t = q.t;
            for (int i = 0; i < t.ni; ++i)
                for (int j = 0; j < t.nj; ++j)
                    for (int k = 0; k < t.nk; ++k) {
                        const int64_t d = t.dst + k + (int64_t)j * t.d_sj + (int64_t)i * t.d_si;
                        const int64_t so = t.src + k + (int64_t)j * t.s_sj + (int64_t)i * t.s_si;
                        const auto f = cur.find(so);
                        const Orig o = f == cur.end() ? Orig{so, q.sg} : f->second;
                        if (cur.insert_or_assign(d, o).second) written.push_back({d, q.dg});
                    }
        }
        std::vector<int64_t> ed, eo;
        std::vector<int> eg, esg;
        std::unordered_map<int64_t, int> at;
        for (const auto &w : written) {
            const Orig o = cur.at(w.first);
            if (o.pos == w.first) continue;   // ends with its own value
            at[w.first] = (int)ed.size();
            ed.push_back(w.first);
            eo.push_back(o.pos);
            eg.push_back(w.second);
            esg.push_back(o.grid);
        }
        const size_t ne = ed.size();
        std::vector<int> nxt(ne, -1), indeg(ne, 0), lev(ne, 0), todo;
        for (size_t e = 0; e < ne; ++e) {
            const auto f = at.find(eo[e]);
            if (f != at.end()) { nxt[e] = f->second; ++indeg[f->second]; }
        }
        for (size_t e = 0; e < ne; ++e)
            if (!indeg[e]) todo.push_back((int)e);
        size_t done = 0;
        int nst = ne ? 1 : 0;
        while (!todo.empty()) {
            const int e = todo.back();
            todo.pop_back();
            ++done;
            const int x = nxt[e];
            if (x < 0) continue;
            lev[x] = std::max(lev[x], lev[e] + 1);
            nst = std::max(nst, lev[x] + 1);
            if (--indeg[x] == 0) todo.push_back(x);
        }
        if (done != ne) throw HipError{"cyclic ghost copies"};
        auto ts = std::make_unique<mgp_handle::TaskSet>();
        ts->valid = valid;
        ts->pairs = false;
        std::vector<PullBlock> pbs;
        std::vector<uint2> ents;
        for (int st = 0; st < nst; ++st) {
            ts->stage_start.push_back((int)pbs.size());
            std::map<int, std::vector<int>> byrecv;
            for (size_t e = 0; e < ne; ++e)
                if (lev[e] == st) byrecv[eg[e]].push_back((int)e);
            for (const auto &kv : byrecv) {
                PullBlock pb{};
                pb.dst = h.ip3[kv.first - 1];
                pb.tab = (int)ents.size();
                std::vector<int> slots;   // source grids
                std::vector<uint2> e2;
                for (const int e : kv.second) {
                    const int sl = (int)(std::find(slots.begin(), slots.end(), esg[e]) - slots.begin());
                    if (sl == (int)slots.size()) {
                        if (sl == 26) throw HipError{"more than 26 ghost copies into one block"};
                        slots.push_back(esg[e]);
                        pb.src[sl] = h.ip3[esg[e] - 1];
                    }
                    const int64_t d = ed[e] - pb.dst, so = eo[e] - pb.src[sl];
                    if (so < 0 || so >= (1 << kPullSlotBits) || d < 0 || d > INT32_MAX)
                        throw HipError{"ghost copy offsets beyond the entry format"};
                    e2.push_back(make_uint2((uint32_t)d, (uint32_t)so | (uint32_t)sl << kPullSlotBits));
                }
                std::sort(e2.begin(), e2.end(), [](const uint2 &a, const uint2 &b) { return a.x < b.x; });
                ents.insert(ents.end(), e2.begin(), e2.end());
                pb.n = (int)e2.size();
                pbs.push_back(pb);
            }
        }
        ts->stage_start.push_back((int)pbs.size());
        upload(ts->blocks, pbs);
        upload(ts->ents, ents);
        auto &ref = *ts;
        h.tasks[key] = std::move(ts);
        return ref;
    }
    for (auto &w : wr) w.clear();
    // staging
    std::vector<int> stage(list.size(), 0);
    int nst = list.empty() ? 0 : 1;
    for (size_t q = 0; q < list.size(); ++q) {
        int st = 0;
        for (int u : wr[list[q].sg])
            if (overlap(list[u].db, list[q].sb)) st = std::max(st, stage[u] + 1);
        for (int u : rdl[list[q].dg])
            if (overlap(list[u].sb, list[q].db)) st = std::max(st, stage[u] + 1);
        for (int u : wr[list[q].dg])
            if (overlap(list[u].db, list[q].db)) st = std::max(st, stage[u] + 1);
        stage[q] = st;
        nst = std::max(nst, st + 1);
        wr[list[q].dg].push_back((int)q);
        rdl[list[q].sg].push_back((int)q);
    }
    // entry tables: per stage and receiving block, the box copies become
    // slots and their elements entries sorted by destination offset (no two
    // copies of a stage write the same position).  Pairs of values along k
    // when every box has an even k extent and even offsets and strides.
    auto ts = std::make_unique<mgp_handle::TaskSet>();
    ts->valid = valid;
    bool pairs = true;
    for (const T2 &q : list) {
        const CopyTask &t = q.t;
        pairs = pairs && t.nk % 2 == 0 && t.src % 2 == 0 && t.dst % 2 == 0 && t.s_sj % 2 == 0 &&
                t.s_si % 2 == 0 && t.d_sj % 2 == 0 && t.d_si % 2 == 0;
    }
    ts->pairs = pairs;
    const int step = pairs ? 2 : 1;
    std::vector<PullBlock> pbs;
    std::vector<uint2> ents;
    std::map<std::vector<int64_t>, std::pair<int, int>> tables;  // pattern -> (tab, n)
    for (int st = 0; st < nst; ++st) {
        ts->stage_start.push_back((int)pbs.size());
        std::map<int, std::vector<int>> byrecv;
        for (size_t q = 0; q < list.size(); ++q)
            if (stage[q] == st) byrecv[list[q].dg].push_back((int)q);
        // receiving blocks in nested neighbour chains (z innermost):
        // consecutive workgroups then work on z-neighbours (the thin
        // z-ghost rows) and neighbouring columns on y-neighbours, so that
        // partial-line reads and writes of the same lines meet in the
        // memory-side cache: 1.5x faster than block-id order on a 32^3
        // lattice numbered x fastest.
        std::vector<int> recvs;
        for (auto &kv : byrecv) recvs.push_back(kv.first);
        const std::vector<int> order = chain_order(h, recvs);
        for (int recv : order) {
            const std::vector<int> &qs = byrecv.at(recv);
            if (qs.size() > 26) throw HipError{"more than 26 ghost copies into one block"};
            PullBlock pb{};
            pb.dst = h.ip3[recv - 1];
            std::vector<int64_t> pat;
            for (size_t sl = 0; sl < qs.size(); ++sl) {
                const T2 &q = list[qs[sl]];
                const CopyTask &t = q.t;
                pb.src[sl] = h.ip3[q.sg - 1];
                pat.insert(pat.end(), {t.dst - pb.dst, t.src - pb.src[sl], t.ni, t.nj, t.nk,
                                       t.s_sj, t.s_si, t.d_sj, t.d_si});
            }
            auto it = tables.find(pat);
            if (it == tables.end()) {
                std::vector<uint2> e;
                for (size_t sl = 0; sl < qs.size(); ++sl) {
                    const CopyTask &t = list[qs[sl]].t;
                    const int64_t d0 = t.dst - pb.dst, s0 = t.src - pb.src[sl];
                    for (int i = 0; i < t.ni; ++i)
                        for (int j = 0; j < t.nj; ++j)
                            for (int k = 0; k < t.nk; k += step) {
                                const int64_t d = d0 + k + (int64_t)j * t.d_sj + (int64_t)i * t.d_si;
                                const int64_t so = s0 + k + (int64_t)j * t.s_sj + (int64_t)i * t.s_si;
                                if (so < 0 || so >= (1 << kPullSlotBits) || d < 0 || d > INT32_MAX)
                                    throw HipError{"ghost copy offsets beyond the entry format"};
                                e.push_back(make_uint2((uint32_t)d, (uint32_t)so | (uint32_t)sl << kPullSlotBits));
                            }
                }
                std::sort(e.begin(), e.end(), [](const uint2 &a, const uint2 &b) { return a.x < b.x; });
                it = tables.emplace(pat, std::make_pair((int)ents.size(), (int)e.size())).first;
                ents.insert(ents.end(), e.begin(), e.end());
            }
            pb.tab = it->second.first;
            pb.n = it->second.second;
            pbs.push_back(pb);
        }
    }
    ts->stage_start.push_back((int)pbs.size());
    upload(ts->blocks, pbs);
    upload(ts->ents, ents);
    auto &ref = *ts;
    h.tasks[key] = std::move(ts);
    return ref;
}

void exchange(mgp_handle &h, int level, int layers, int corners, int forward,
              double *const *fields, int nfield, bool kpush = false) {
    auto &ts = task_set(h, level, layers, corners, forward, kpush);
    for (int f0 = 0; f0 < nfield; f0 += 6) {
        const int nf = std::min(6, nfield - f0);
        int mode = 0;
        if (ts.pairs) {
            mode = 2;
            for (int f = 0; f < nf; ++f)
                if ((uintptr_t)fields[f0 + f] % 16) mode = 1;
        }
        for (size_t s = 0; s + 1 < ts.stage_start.size(); ++s) {
            const int a = ts.stage_start[s], b = ts.stage_start[s + 1];
            launch_pull(ts.blocks.as<PullBlock>() + a, b - a, ts.ents.as<uint2>(), mode,
                        fields + f0, nf, h.stream);
        }
    }
}

// LDS per workgroup that the factorisation and the plane kernels need for
// the handle's largest block (the sums of their launchers, the HILF window
// at its four-slot maximum).
size_t plane_lds_need(const mgp_handle &h) {
    const size_t T = h.tsz();
    const size_t mRS = (size_t)(h.max_nj + 1) * (h.max_ni + 1);
    const size_t mHS = (size_t)(h.max_nj + 2) * (h.max_ni + 2);
    const int m = std::max(std::max(h.max_ni, h.max_nj), h.max_nk);
    const size_t nr = (size_t)((m + 1) & ~1);
    const size_t tab = 24 * sizeof(void *);   // source table: three rows of 8 pointers
    const size_t mb = h.masked ? mHS : 0;
    const size_t maxc = (size_t)h.max_ni * h.max_nj * h.max_nk;
    size_t need = std::min(maxc, (size_t)kStage) * sizeof(double);                    // divcal
    need = std::max(need, 6 * mRS * sizeof(double));                                 // k_lu
    need = std::max(need, tab + 6 * nr * T + 2 * mRS * T + 3 * nr * 8 + 4 * (mHS * T + mb));  // forward
    need = std::max(need, tab + 9 * nr * 8 + 2 * mRS * T + 4 * (mHS * T + mb));     // fused outer
    if (h.cfg.ityp == 1)
        need = std::max(need, tab + 6 * nr * T + 8 * mHS * T + (h.masked ? 8 * mHS : 3 * nr * 8));
    return need;
}

// Decided once per handle: whether mgpoisl's velocity and pressure
// correction runs along z-chains and writes the k-face ghosts of interior
// rows for the closing exchange (k_mgpcorr_chain).  It does on single-level
// grids with no PAR face and no face bound_flow acts on (so nothing runs
// between the correction and the exchange), when both k-faces of every
// block are connected to blocks of its size with nk >= 2, the k-high
// neighbour of each block takes its k-low ghosts from that block (so the
// connections form cycles along z), and no copy left in the exchange reads
// or writes a region the correction fills.  MGP_KPUSH=0 in the environment
// keeps the plain path.
bool kpush_on(mgp_handle &h) {
    if (h.kpush_state) return h.kpush_state > 0;
    h.kpush_state = -1;
    const char *env = getenv("MGP_KPUSH");
    if (env && atoi(env) == 0) return false;
    if (h.level_blocks.size() != 1 || !h.level_par.empty() || !h.level_flow.empty())
        return false;
    // up[g]: the block whose k = 2, 3 fill g's k-high ghosts; down[g]: the
    // block whose k = nk, nk + 1 fill g's k-low ghosts (0-based, -1 none)
    std::vector<int> up(h.ngrid, -1), down(h.ngrid, -1);
    for (const Conn &c : h.conns) {
        if (c.frecv != 5 && c.frecv != 6) continue;
        const int d = c.recv - 1, sg = c.send - 1;
        if (h.ii[d] != h.ii[sg] || h.jj[d] != h.jj[sg] || h.kk[d] != h.kk[sg] || h.kk[d] < 6 ||
            h.kk[d] > kCorrThreads)
            return false;
        int &nb = c.frecv == 6 ? up[d] : down[d];
        if (nb >= 0) return false;
        nb = sg;
    }
    for (int g = 0; g < h.ngrid; ++g)
        if (up[g] < 0 || down[g] < 0 || down[up[g]] != g) return false;
    h.kpush_faces.assign(h.ngrid, 3);
    if (!task_set(h, h.level_blocks.begin()->first, 2, 1, 0, true).valid) {
        h.kpush_faces.clear();
        return false;
    }
    // cycles along +z, started from the blocks in id order (chains that
    // run at the same time then hold blocks close in memory)
    std::vector<char> seen(h.ngrid, 0);
    std::vector<int> chain;
    std::vector<int2> chains;
    int nx = 0;
    for (int g1 = 1; g1 <= h.ngrid; ++g1) {
        if (seen[g1 - 1]) continue;
        const int start = (int)chain.size();
        for (int g = g1 - 1; !seen[g]; g = up[g]) {
            seen[g] = 1;
            chain.push_back(g);
        }
        if ((int)chain.size() - start > kCorrChainMax) return false;
        chains.push_back(make_int2(start, (int)chain.size() - start));
        const int nrow = (h.ii[g1 - 1] - 2) * (h.jj[g1 - 1] - 2);
        const int rows = std::min(kCorrThreads / h.kk[g1 - 1], kCorrRows);
        nx = std::max(nx, (nrow + rows - 1) / rows);
    }
    upload(h.d_cchain, chain);
    upload(h.d_cchains, chains);
    h.ncchain = (int)chains.size();
    h.ccnx = nx;
    h.kpush_state = 1;
    return true;
}

// ------------------------------------------------------------------
// Plans of the plane kernels (mgp_sweeps.hip), per shape.  Step p of a
// plan computes hyperplane p (SOR: the plane its colour schedule relaxes)
// and moves the entries that later steps read into LDS: the HILF window
// plane p+2 into ring slot (p+3) mod 4 (mod 8 for SOR), and the ghosts of
// the recurrence that step p+1 reads (the backward sweep: p-1) into pencil
// buffer p mod 2.  Steps before the first plane only fill LDS.
void build_plans(mgp_handle &h, const std::vector<uint32_t> &cells,
                 const std::vector<int> &pstart, std::vector<uint2> &items,
                 std::vector<uint32_t> &ents, std::vector<uint32_t> &cellx,
                 std::vector<uint32_t> &fents) {
    cellx.resize(cells.size());
    for (Shape &S : h.shapes) {
        const int NI = S.ni, NJ = S.nj, NK = S.nk, NP = S.np;
        const int HW = NJ + 2, RW = NJ + 1;
        const uint32_t *cl = cells.data() + S.cell_off;
        const int *ps = pstart.data() + S.pstart_off;
        // cell words: indices and boundary flags
        for (int s = 0; s < S.ncell; ++s) {
            const int i = cl[s] & 1023, j = (cl[s] >> 10) & 1023, k = cl[s] >> 20;
            const uint32_t fl = (i == 0) | (i == NI - 1) << 1 | (j == 0) << 2 |
                                (j == NJ - 1) << 3 | (k == 0) << 4 | (k == NK - 1) << 5;
            cellx[S.cell_off + s] = (uint32_t)i | (uint32_t)j << 6 | (uint32_t)k << 12 | fl << 18;
        }
        using Ents = std::vector<uint32_t>;
        // HILF window plane q: its interior cells, then the face ghosts
        // (i, j, -1), (i, j, NK), (i, -1, k), (i, NJ, k), (-1, j, k),
        // (NI, j, k) that lie on plane q, each at its pencil position
        auto window = [&](int q, Ents &E) {
            if (q >= 0 && q < NP)
                for (int s = ps[q]; s < ps[q + 1]; ++s) {
                    const int i = cl[s] & 1023, j = (cl[s] >> 10) & 1023;
                    E.push_back(plan_entry((j + 1) + HW * (i + 1), 6, 0, s));
                }
            for (int i = 0; i < NI; ++i) {
                const int j = q + 1 - i;
                if (j >= 0 && j < NJ) E.push_back(plan_entry((j + 1) + HW * (i + 1), 4, 0, fidx_z(j, i, NI, NJ)));
            }
            for (int i = 0; i < NI; ++i) {
                const int j = q - NK - i;
                if (j >= 0 && j < NJ) E.push_back(plan_entry((j + 1) + HW * (i + 1), 5, 0, fidx_z(j, i, NI, NJ)));
            }
            for (int i = 0; i < NI; ++i) {
                const int k = q + 1 - i;
                if (k >= 0 && k < NK) E.push_back(plan_entry(HW * (i + 1), 2, 0, fidx_y(k, i, NI, NK)));
            }
            for (int i = 0; i < NI; ++i) {
                const int k = q - NJ - i;
                if (k >= 0 && k < NK) E.push_back(plan_entry((NJ + 1) + HW * (i + 1), 3, 0, fidx_y(k, i, NI, NK)));
            }
            for (int j = 0; j < NJ; ++j) {
                const int k = q + 1 - j;
                if (k >= 0 && k < NK) E.push_back(plan_entry(j + 1, 0, 0, fidx_x(k, j, NJ, NK)));
            }
            for (int j = 0; j < NJ; ++j) {
                const int k = q - NI - j;
                if (k >= 0 && k < NK) E.push_back(plan_entry((j + 1) + HW * (NI + 1), 1, 0, fidx_x(k, j, NJ, NK)));
            }
        };
        // forward recurrence: pencil starts (i, j, -1), and the ghosts
        // (i, -1, k), (-1, j, k) read by the cells of plane pp
        auto fwd_start = [&](Ents &E) {
            for (int i = 0; i < NI; ++i)
                for (int j = 0; j < NJ; ++j)
                    E.push_back(plan_entry((j + 1) + RW * (i + 1), 4, 1, fidx_z(j, i, NI, NJ)));
        };
        auto fwd_border = [&](int pp, Ents &E) {
            for (int i = 0; i < NI; ++i) {
                const int k = pp - i;
                if (k >= 0 && k < NK) E.push_back(plan_entry(RW * (i + 1), 2, 1, fidx_y(k, i, NI, NK)));
            }
            for (int j = 0; j < NJ; ++j) {
                const int k = pp - j;
                if (k >= 0 && k < NK) E.push_back(plan_entry(j + 1, 0, 1, fidx_x(k, j, NJ, NK)));
            }
        };
        // backward recurrence: pencil starts (i, j, NK), ghosts (i, NJ, k),
        // (NI, j, k) read by the cells of plane pp
        auto bwd_start = [&](Ents &E) {
            for (int i = 0; i < NI; ++i)
                for (int j = 0; j < NJ; ++j)
                    E.push_back(plan_entry(j + RW * i, 5, 1, fidx_z(j, i, NI, NJ)));
        };
        auto bwd_border = [&](int pp, Ents &E) {
            for (int i = 0; i < NI; ++i) {
                const int k = pp - (NJ - 1) - i;
                if (k >= 0 && k < NK) E.push_back(plan_entry(NJ + RW * i, 3, 1, fidx_y(k, i, NI, NK)));
            }
            for (int j = 0; j < NJ; ++j) {
                const int k = pp - j - (NI - 1);
                if (k >= 0 && k < NK) E.push_back(plan_entry(j + RW * NI, 1, 1, fidx_x(k, j, NJ, NK)));
            }
        };
        auto plane = [&](int q, int &cpos, int &ccnt) {
            if (q >= 0 && q < NP) { cpos = ps[q]; ccnt = ps[q + 1] - ps[q]; }
            else { cpos = 0; ccnt = 0; }
        };
        for (int kind = 0; kind < kPlanKinds; ++kind) {
            S.plan_ent[kind] = (int)ents.size();
            const int ebase0 = (int)ents.size();
            struct Step { int p, cpos, ccnt, epos, ecnt; };
            std::vector<Step> steps;
            std::vector<int> order;
            if (kind == kPlanBwd)
                for (int p = NP + 1; p >= 0; --p) order.push_back(p);
            else
                for (int p = -3; p < (kind == kPlanSor ? NP + 2 : NP); ++p) order.push_back(p);
            for (int p : order) {
                Ents E;
                int cpos, ccnt;
                switch (kind) {
                case kPlanFwdLap:
                case kPlanFwdZero:
                    if (kind == kPlanFwdLap && p + 2 <= NP) window(p + 2, E);
                    if (p == -2 || p == -1) fwd_start(E);
                    if (p >= -1 && p + 1 < NP) fwd_border(p + 1, E);
                    plane(p, cpos, ccnt);
                    break;
                case kPlanBwd:
                    if (p >= NP) bwd_start(E);
                    if (p <= NP && p >= 1) bwd_border(p - 1, E);
                    plane(p, cpos, ccnt);
                    break;
                case kPlanOut:
                    if (p + 2 <= NP) window(p + 2, E);
                    plane(p, cpos, ccnt);
                    break;
                default:  // kPlanSor
                    if (p + 2 <= NP) window(p + 2, E);
                    plane(p < 0 ? -1 : ((p & 1) ? p : p - 2), cpos, ccnt);
                    break;
                }
                steps.push_back({p, cpos, ccnt, (int)ents.size() - ebase0, (int)E.size()});
                ents.insert(ents.end(), E.begin(), E.end());
            }
            // one valid entry for the clamped loads of empty items
            ents.push_back(plan_entry(0, 7, 0, 0));
            auto pack = [](int cpos, int cc, int ec, int epos, int p) {
                uint2 v;
                v.x = (uint32_t)cpos | (uint32_t)cc << 16 | (uint32_t)ec << 24;
                v.y = (uint32_t)epos | (uint32_t)(p + 8) << 21;
                return v;
            };
            // One wave per block keeps three window slots in the FP64 kernels
            // that read HILF's stencil (win_slots): an entry of window(p+2) lands in
            // the slot plane p-1 occupies, so it goes to an item no earlier
            // than the one computing the last cell of plane p that reads
            // plane p-1 at its position (the cell itself, its j+1 and its i+1
            // neighbour: Hm[c0], Hm[c0-1], Hm[c0-HW]); entries are reordered
            // per step by that item and spill into entry-only items of the
            // same step where the step's compute items cannot take them.
            const bool three = (kind == kPlanFwdLap || kind == kPlanOut) && !h.fp32();
            const int HS = (NI + 2) * HW;
            // cell chunks (first, count) of a step: at most W cells and
            // kFaceSlots face entries each (a chunk ends early only where
            // a block is one cell thick along an axis and all its cells
            // lie on two faces)
            auto chunks = [&](const Step &st, int W) {
                std::vector<std::pair<int, int>> ch;
                for (int t = 0; t < st.ccnt;) {
                    int n = 0, f = 0;
                    while (t + n < st.ccnt && n < W) {
                        const uint32_t w = cl[st.cpos + t + n];
                        const int i = w & 1023, j = (w >> 10) & 1023, k = w >> 20;
                        const int nf = (i == 0) + (i == NI - 1) + (j == 0) + (j == NJ - 1) +
                                       (k == 0) + (k == NK - 1);
                        if (f + nf > kFaceSlots) break;
                        f += nf;
                        ++n;
                    }
                    ch.push_back({t, n});
                    t += n;
                }
                return ch;
            };
            auto scheduled = [&](int W) {
                std::vector<int> cellitem(HS);
                for (const Step &st : steps) {
                    std::fill(cellitem.begin(), cellitem.end(), -1);
                    const auto ch = chunks(st, W);
                    const int nc = (int)ch.size();
                    for (int c = 0; c < nc; ++c)
                        for (int t = ch[c].first; t < ch[c].first + ch[c].second; ++t) {
                            const uint32_t w = cl[st.cpos + t];
                            cellitem[((w >> 10) & 1023) + 1 + HW * ((w & 1023) + 1)] = c;
                        }
                    std::vector<std::pair<int, uint32_t>> es;
                    for (int q = 0; q < st.ecnt; ++q) {
                        const uint32_t e = ents[ebase0 + st.epos + q];
                        int last = 0;
                        if (!(e & 0x8000u)) {
                            const int x = e & 0xfff;
                            for (const int c : {x, x + 1, x + HW})
                                if (c < HS && cellitem[c] > last) last = cellitem[c];
                        }
                        es.push_back({last, e});
                    }
                    std::stable_sort(es.begin(), es.end(),
                                     [](const auto &a, const auto &b) { return a.first < b.first; });
                    size_t idx = 0;
                    for (int c = 0; c < nc || idx < es.size(); ++c) {
                        const int epos = (int)ents.size() - ebase0;
                        int ec = 0;
                        while (idx < es.size() && ec < W && es[idx].first <= c) {
                            ents.push_back(es[idx++].second);
                            ++ec;
                        }
                        const int cc = c < nc ? ch[c].second : 0;
                        items.push_back(pack(cc ? st.cpos + ch[c].first : 0, cc, ec, ec ? epos : 0, st.p));
                    }
                }
            };
            for (int nwi = 0; nwi < 4; ++nwi) {
                const int NW = kPlanWaves[nwi];
                for (int wi = 0; wi < 2; ++wi) {
                    const int W = 64 << wi;
                    const size_t base = items.size();
                    S.plan_item[nwi][wi][kind] = (int)base;
                    if (three && NW == 1) scheduled(W);
                    else for (const Step &st : steps) {
                        const auto ch = chunks(st, W);
                        const int nc = (int)ch.size();
                        const int n = std::max(nc, (st.ecnt + W - 1) / W);
                        const int rows = (n + NW - 1) / NW;
                        for (int c = 0; c < rows * NW; ++c) {
                            if (c >= n) {   // padding: a no-op of the same step
                                items.push_back(pack(0, 0, 0, 0, st.p));
                                continue;
                            }
                            const int cc = c < nc ? ch[c].second : 0;
                            const int ec = std::min(std::max(st.ecnt - W * c, 0), W);
                            items.push_back(pack(cc ? st.cpos + ch[c].first : 0, cc, ec,
                                                 ec ? st.epos + W * c : 0, st.p));
                        }
                    }
                    if ((int)ents.size() - ebase0 >= (1 << 21))
                        throw HipError{"plan entry table too large"};
                    // closing no-op items (the engine clamps item indices to
                    // the last one)
                    do
                        for (int w = 0; w < NW; ++w) items.push_back(pack(0, 0, 0, 0, 0));
                    while (((items.size() - base) / NW) % kPlanPad);
                    const int per = (int)((items.size() - base) / NW);
                    S.plan_nitem[nwi][wi][kind] = per;
                    // every wave must meet the same steps (its barriers)
                    for (int q = 0; q < per; ++q)
                        for (int w = 1; w < NW; ++w)
                            if ((items[base + (size_t)q * NW + w].y >> 21) !=
                                (items[base + (size_t)q * NW].y >> 21))
                                throw HipError{"plan steps differ between waves"};
                }
            }
        }
    }
    // face-store entries, 64 per item (parallel to the item table): the
    // boundary cells of the item's compute chunk, one entry per (cell, own
    // face) -- the forward sweeps' RES from the recurrence plane (lower
    // faces, and upper faces flagged), the backward sweep's HILF from the
    // item's staging row, SOR's new value from its window slot
    fents.assign(items.size() * kFaceSlots, kFaceNone);
    for (const Shape &S : h.shapes) {
        const int NI = S.ni, NJ = S.nj, NK = S.nk;
        const uint32_t *cl = cells.data() + S.cell_off;
        for (int kind = 0; kind < kPlanKinds; ++kind) {
            if (kind == kPlanOut) continue;
            for (int nwi = 0; nwi < 4; ++nwi)
                for (int wi = 0; wi < 2; ++wi) {
                    const size_t base = S.plan_item[nwi][wi][kind];
                    const size_t n = (size_t)S.plan_nitem[nwi][wi][kind] * kPlanWaves[nwi];
                    for (size_t q = base; q < base + n; ++q) {
                        const int cpos = items[q].x & 0xffff, cc = (items[q].x >> 16) & 0xff;
                        int e = 0;
                        for (int l = 0; l < cc; ++l) {
                            const uint32_t w = cl[cpos + l];
                            const int i = w & 1023, j = (w >> 10) & 1023, k = w >> 20;
                            const bool on[6] = {i == 0, i == NI - 1, j == 0, j == NJ - 1,
                                                k == 0, k == NK - 1};
                            const int pos[3] = {fidx_x(k, j, NJ, NK), fidx_y(k, i, NI, NK),
                                                fidx_z(j, i, NI, NJ)};
                            const int src = kind == kPlanBwd ? l
                                          : kind == kPlanSor ? (j + 1) + (NJ + 2) * (i + 1)
                                                             : (j + 1) + (NJ + 1) * (i + 1);
                            for (int f = 0; f < 6; ++f) {
                                if (!on[f]) continue;
                                if (e == kFaceSlots)
                                    throw HipError{"more face cells in a plan item than lanes"};
                                fents[q * kFaceSlots + e++] =
                                    face_entry(src, face_base(f, NI, NJ, NK) + pos[f >> 1], f & 1);
                            }
                        }
                    }
                }
        }
    }
}

// ------------------------------------------------------------------
// Shape tables: hyperplane order of src/flow/sip_hyperplane_mod.F90:
// 269-361 (planes m = i+j+k, ascending linear index inside a plane, i.e.
// i ascending then j ascending).
void build_shapes(mgp_handle &h, std::vector<uint32_t> &cells,
                  std::vector<int> &pstart, std::vector<int> &slot) {
    std::map<std::tuple<int, int, int>, int> idx;
    h.blocks.resize(h.ngrid);
    int64_t wo = 0, fc = 0, sh = 0;
    for (int g = 0; g < h.ngrid; ++g) {
        const int ni = h.ii[g] - 4, nj = h.jj[g] - 4, nk = h.kk[g] - 4;
        auto key = std::make_tuple(ni, nj, nk);
        auto it = idx.find(key);
        int sid;
        if (it == idx.end()) {
            Shape S;
            S.ni = ni; S.nj = nj; S.nk = nk;
            S.np = ni + nj + nk - 2;
            S.ncell = ni * nj * nk;
            S.cell_off = (int64_t)cells.size();
            S.pstart_off = (int64_t)pstart.size();
            S.slot_off = (int64_t)slot.size();
            S.fsize = 2 * (nk * nj + nk * ni + nj * ni);
            slot.resize(slot.size() + S.ncell);
            int s = 0;
            for (int p = 0; p < S.np; ++p) {
                pstart.push_back(s);
                for (int i = 0; i < ni; ++i)
                    for (int j = 0; j < nj; ++j) {
                        const int k = p - i - j;
                        if (k < 0 || k >= nk) continue;
                        cells.push_back((uint32_t)i | ((uint32_t)j << 10) |
                                        ((uint32_t)k << 20));
                        slot[S.slot_off + k + nk * (j + nj * i)] = s;
                        ++s;
                    }
            }
            pstart.push_back(s);
            // face_index must be a bijection onto [0, a*b) for each face
            const int fa[3][2] = {{nk, nj}, {nk, ni}, {nj, ni}};
            for (auto &ab : fa) {
                std::vector<char> seen((size_t)ab[0] * ab[1], 0);
                for (int u = 0; u < ab[0]; ++u)
                    for (int v = 0; v < ab[1]; ++v) {
                        const int q = face_index(u, v, ab[0], ab[1]);
                        if (q < 0 || q >= ab[0] * ab[1] || seen[q]++)
                            throw HipError{"face index table inconsistent"};
                    }
            }
            for (int p = 0; p < S.np; ++p)
                h.max_plane = std::max(h.max_plane, pstart[S.pstart_off + p + 1] - pstart[S.pstart_off + p]);
            sid = (int)h.shapes.size();
            h.shapes.push_back(S);
            idx[key] = sid;
        } else {
            sid = it->second;
        }
        Block &B = h.blocks[g];
        B.shape = sid;
        B.ni = ni; B.nj = nj; B.nk = nk;
        B.level = h.level[g];
        B.flags = 0;
        B.par = 0;
        B.parent = h.iparent.empty() ? -1 : h.iparent[g] - 1;
        for (int a = 0; a < 3; ++a) B.pos[a] = h.position.empty() ? 0 : h.position[3 * g + a];
        for (int f = 0; f < 6; ++f) {
            B.bc[f] = h.bc[6 * g + f];
            const int ng = h.nbr[26 * g + f];
            B.nbr[f] = (is_con(B.bc[f]) && ng > 0) ? ng - 1 : -1;
            if (B.bc[f] == PAR_BC) B.par |= 1 << f;
        }
        B.wo = wo;
        B.fo = wo;
        B.sp = h.ip3[g];
        B.fc = fc;
        B.sh = sh;
        B.o1x = h.ip1x[g]; B.o1y = h.ip1y[g]; B.o1z = h.ip1z[g];
        wo += (int64_t)ni * nj * nk;
        fc += h.shapes[sid].fsize;
        sh += shell_size(ni, nj, nk);
        h.max_ni = std::max(h.max_ni, ni);
        h.max_nj = std::max(h.max_nj, nj);
        h.max_nk = std::max(h.max_nk, nk);
    }
    h.ncell = wo;
    h.nface = fc;
    h.nshell = sh;
}

template <typename T>
void upload(DBuf &d, const std::vector<T> &v) {
    d.alloc(v.size() * sizeof(T));
    if (!v.empty())
        HC(hipMemcpy(d.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}

void upload_raw(DBuf &d, const void *src, size_t bytes) {
    d.alloc(bytes);
    if (bytes) HC(hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice));
}

int check_cfg(const mgp_config &c) {
    if (c.ityp < 0 || c.ityp > 2) return fail(MGP_ERR_ARG, "invalid pressure solver type %d", c.ityp);
    if (c.ninner < 0) return fail(MGP_ERR_ARG, "Invalid number of ninner: %d", c.ninner);
    if (c.nouter < 0) return fail(MGP_ERR_ARG, "Invalid number of nouter: %d", c.nouter);
    if (c.nouter_min < 0) return fail(MGP_ERR_ARG, "Invalid number of nouter_min: %d", c.nouter_min);
    if (c.epcorr < 0.0) return fail(MGP_ERR_ARG, "Invalid epcorr: %g", c.epcorr);
    if (c.ityp == 1 && (c.omega <= 0.0 || c.omega >= 2.0))
        return fail(MGP_ERR_ARG, "Invalid omega: %g", c.omega);
    if (c.precision < 0 || c.precision > 1) return fail(MGP_ERR_ARG, "invalid precision %d", c.precision);
    if (c.ibtype < 0 || c.ibtype > 1) return fail(MGP_ERR_ARG, "invalid ibtype %d", c.ibtype);
    if (!(c.rho > 0.0)) return fail(MGP_ERR_ARG, "invalid rho %g", c.rho);
    if (!(c.alfa <= 1.0)) return fail(MGP_ERR_ARG, "invalid SIP alfa %g", c.alfa);
    if (c.sync_by_caller < 0 || c.sync_by_caller > 1)
        return fail(MGP_ERR_ARG, "invalid sync_by_caller %d", c.sync_by_caller);
    return MGP_OK;
}

// ------------------------------------------------------------------
template <typename T>
void setup_factors(mgp_handle &h) {
    Topo t = h.topo(-1);
    T *f[7];
    for (int q = 0; q < 7; ++q) f[q] = h.d_f[q].as<T>();
    launch_lu<T>(t, h.d_ap64.as<double>(), h.d_c64[0].as<double>(),
                 h.d_c64[1].as<double>(), h.d_c64[2].as<double>(),
                 h.d_c64[3].as<double>(), h.d_c64[4].as<double>(),
                 h.d_c64[5].as<double>(), h.mask(), h.cfg.alfa, f, h.stream);
    HC(hipGetLastError());
}

// Shared factor storage: groups the blocks whose SIP factors, GSAP (and BP
// bytes) are equal bit for bit, keeps one copy of the factors and GSAP per
// group and points each block's factor offset (Block::fo) at it.  The
// values are the same, so every result is unchanged; the plane kernels read
// the shared copy from L2 instead of per-block copies from HBM.
// MGP_SHARED_FACTORS=0 keeps one copy per block.
template <typename T>
void share_factors(mgp_handle &h) {
    const hipStream_t s = h.stream;
    const int n = h.ngrid;
    h.frep.resize(n);
    for (int b = 0; b < n; ++b) h.frep[b] = b;
    h.nfclass = n;
    h.nfcell = h.ncell;
    const char *env = getenv("MGP_SHARED_FACTORS");
    const bool on = !(env && atoi(env) == 0);
    Topo t = h.topo(-1);
    const T *f[7];
    for (int q = 0; q < 7; ++q) f[q] = h.d_f[q].as<T>();
    Mask m = h.mask();
    const double *ap = h.d_ap64.as<double>();
    if (on && n > 1) {
        DBuf dh, drep, dbad;
        dh.alloc((size_t)n * 8);
        launch_fhash<T>(t, f, ap, m, dh.as<unsigned long long>(), s);
        HC(hipGetLastError());
        std::vector<unsigned long long> hash(n);
        HC(hipMemcpyAsync(hash.data(), dh.p, (size_t)n * 8, hipMemcpyDeviceToHost, s));
        HC(hipStreamSynchronize(s));
        std::map<std::pair<unsigned long long, int>, int> first;
        for (int b = 0; b < n; ++b) {
            auto key = std::make_pair(hash[b], h.blocks[b].shape);
            auto it = first.find(key);
            if (it == first.end()) first[key] = b;
            else h.frep[b] = it->second;
        }
        // confirm every member against its representative, bit for bit
        upload(drep, h.frep);
        std::vector<int> bad(n, 0);
        upload(dbad, bad);
        launch_fcmp<T>(t, f, ap, m, drep.as<int>(), dbad.as<int>(), s);
        HC(hipGetLastError());
        HC(hipMemcpyAsync(bad.data(), dbad.p, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, s));
        HC(hipStreamSynchronize(s));
        for (int b = 0; b < n; ++b)
            if (bad[b]) h.frep[b] = b;      // hash collision: a class of its own
    }
    std::vector<int> reps;
    for (int b = 0; b < n; ++b)
        if (h.frep[b] == b) reps.push_back(b);
    upload(h.d_blist_rep, reps);
    h.nfclass = (int)reps.size();
    if (h.nfclass == n) return;         // nothing shared: fo stays wo
    // new arrays: the representatives' factors, class after class
    int64_t off = 0;
    for (int r : reps) {
        h.blocks[r].fo = off;
        off += h.shapes[h.blocks[r].shape].ncell;
    }
    for (int b = 0; b < n; ++b) h.blocks[b].fo = h.blocks[h.frep[b]].fo;
    h.nfcell = off;
    DBuf nf[7];
    T *g[7];
    for (int q = 0; q < 7; ++q) {
        nf[q].alloc((size_t)off * sizeof(T));
        g[q] = nf[q].as<T>();
    }
    h.d_aps64.alloc((size_t)off * sizeof(double));
    if (h.fp32()) h.d_apsT.alloc((size_t)off * sizeof(float));
    // the gather reads Block::wo and writes Block::fo of the new blocks table
    upload(h.d_blocks, h.blocks);
    launch_fgather<T>(h.topo_reps(), f, ap, g, h.d_aps64.as<double>(),
                      h.fp32() ? h.d_apsT.as<float>() : nullptr, s);
    HC(hipGetLastError());
    HC(hipStreamSynchronize(s));
    for (int q = 0; q < 7; ++q) {   // the old arrays go with nf
        std::swap(h.d_f[q].p, nf[q].p);
        std::swap(h.d_f[q].n, nf[q].n);
    }
}

template <typename T>
void tune_shape(mgp_handle &h);

int init_common(const mgp_grids *G, const mgp_config *cfg, const double *dx,
                const double *dy, const double *dz, const double *ddx,
                const double *ddy, const double *ddz, const double *bp,
                const int8_t *blocked, mgp_handle **out) {
    if (!G || !cfg || !out || !dx || !dy || !dz || !ddx || !ddy || !ddz)
        return fail(MGP_ERR_ARG, "null argument");
    *out = nullptr;
    int rc = check_cfg(*cfg);
    if (rc) return rc;
    if (G->ngrid <= 0) return fail(MGP_ERR_TOPOLOGY, "no grids");
    auto h = std::make_unique<mgp_handle>();
    h->cfg = *cfg;
    if (h->cfg.check_every < 1) h->cfg.check_every = 1;
    // alfa 0 is the reference's default-real literal 0.92 in an FP64 build
    if (h->cfg.alfa == 0.0) h->cfg.alfa = (double)0.92f;
    else if (h->cfg.alfa < 0.0) h->cfg.alfa = 0.0;
    h->ngrid = G->ngrid;
    h->minlevel = G->minlevel;
    h->maxlevel = G->maxlevel;
    const int n = G->ngrid;
    h->kk.assign(G->kk, G->kk + n);
    h->jj.assign(G->jj, G->jj + n);
    h->ii.assign(G->ii, G->ii + n);
    h->level.assign(G->level, G->level + n);
    h->nbr.assign(G->nbrgrid, G->nbrgrid + 26 * n);
    h->bc.assign(G->itypbc, G->itypbc + 6 * n);
    if (G->iparent) {
        if (!G->position) return fail(MGP_ERR_ARG, "iparent given without position");
        h->iparent.assign(G->iparent, G->iparent + n);
        h->position.assign(G->position, G->position + 3 * n);
    }
    for (int g = 0; g < n; ++g) {
        h->ip3.push_back(G->ip3[g] - 1);
        h->ip1x.push_back(G->ip1x[g] - 1);
        h->ip1y.push_back(G->ip1y[g] - 1);
        h->ip1z.push_back(G->ip1z[g] - 1);
        const int ni = h->ii[g] - 4, nj = h->jj[g] - 4, nk = h->kk[g] - 4;
        // plane kernels: 1-D coefficient rows of kMaxN1 entries, 16-bit
        // cell positions in plan items and entries, 12-bit LDS positions in
        // the pencil and window planes of (ni+2)(nj+2) cells, 14-bit
        // offsets into the block's face region (the LDS the kernels need
        // is checked against the device once the shapes are known)
        if (ni < 1 || nj < 1 || nk < 1 || ni > kMaxN1 || nj > kMaxN1 || nk > kMaxN1 ||
            ni * nj * nk > 65536 || (ni + 2) * (nj + 2) > 4096 ||
            2 * (nk * nj + nk * ni + nj * ni) > 16384)
            return fail(MGP_ERR_TOPOLOGY, "grid %d: interior %dx%dx%d outside the supported "
                        "block sizes (each 1..%d, at most 65536 cells)", g + 1, ni, nj, nk,
                        kMaxN1);
        if (h->level[g] < h->minlevel || h->level[g] > h->maxlevel)
            return fail(MGP_ERR_TOPOLOGY, "grid %d: level out of range", g + 1);
        if (g > 0 && h->level[g] < h->level[g - 1])
            return fail(MGP_ERR_TOPOLOGY, "grids must be numbered level by level");
        // nesting (src/ib/ctof/ctof1_mod.F90:240-255, parent_core_mod.F90:
        // 214-262): a grid above the coarsest level has a parent one level
        // coarser, even cell counts, and covers parent cells only
        const int par = h->iparent.empty() ? 0 : h->iparent[g];
        bool has_par_face = false;
        for (int f = 0; f < 6; ++f) has_par_face = has_par_face || h->bc[6 * g + f] == PAR_BC;
        if (h->level[g] > h->minlevel || par != 0 || has_par_face) {
            if (par < 1 || par > n || h->level[par - 1] != h->level[g] - 1)
                return fail(MGP_ERR_TOPOLOGY, "grid %d: level %d needs a parent on level %d",
                            g + 1, h->level[g], h->level[g] - 1);
            if ((ni | nj | nk) & 1)
                return fail(MGP_ERR_TOPOLOGY, "grid %d: refined grids need even cell counts", g + 1);
            const int pn[3] = {h->ii[par - 1] - 4, h->jj[par - 1] - 4, h->kk[par - 1] - 4};
            const int cn[3] = {ni, nj, nk};
            for (int a = 0; a < 3; ++a) {
                const int p0 = h->position[3 * g + a];
                if (p0 < 3 || p0 - 3 + cn[a] / 2 > pn[a])
                    return fail(MGP_ERR_TOPOLOGY, "grid %d: position outside its parent", g + 1);
            }
        }
        h->level_blocks[h->level[g]].push_back(g);
        h->n3d = std::max<int64_t>(h->n3d, h->ip3[g] + (int64_t)h->ii[g] * h->jj[g] * h->kk[g]);
        h->n1x = std::max<int64_t>(h->n1x, h->ip1x[g] + h->ii[g]);
        h->n1y = std::max<int64_t>(h->n1y, h->ip1y[g] + h->jj[g]);
        h->n1z = std::max<int64_t>(h->n1z, h->ip1z[g] + h->kk[g]);
    }
    // a CON face must see a neighbour of matching face size
    for (int g = 0; g < n; ++g)
        for (int f = 0; f < 6; ++f) {
            const int ng = h->nbr[26 * g + f];
            if (!is_con(h->bc[6 * g + f]) || ng <= 0) continue;
            if (ng > n) return fail(MGP_ERR_TOPOLOGY, "grid %d: neighbour id %d out of range", g + 1, ng);
            const int a = ng - 1;
            const bool ok = (f < 2) ? (h->jj[a] == h->jj[g] && h->kk[a] == h->kk[g])
                          : (f < 4) ? (h->ii[a] == h->ii[g] && h->kk[a] == h->kk[g])
                                    : (h->ii[a] == h->ii[g] && h->jj[a] == h->jj[g]);
            if (!ok || h->level[a] != h->level[g])
                return fail(MGP_ERR_TOPOLOGY, "grid %d face %d: neighbour %d does not match", g + 1, f + 1, ng);
        }
    try {
        HC(hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking));
        h->stream = h->own_stream;
        h->scratch = static_cast<double *>(scratch_device_ptr());
        std::vector<uint32_t> cells;
        std::vector<int> pstart, slot;
        build_shapes(*h, cells, pstart, slot);
        std::vector<uint2> items;
        std::vector<uint32_t> ents;
        std::vector<uint32_t> cellx;
        std::vector<uint32_t> fents;
        build_plans(*h, cells, pstart, items, ents, cellx, fents);
        build_conns(*h);

        // BP: per WO cell, own ghost layer per face, reference layout
        std::vector<double> bpd;
        if (bp || blocked) {
            bpd.resize(h->n3d);
            for (int64_t c = 0; c < h->n3d; ++c)
                bpd[c] = bp ? bp[c] : (blocked[c] ? 0.0 : 1.0);
        }
        std::vector<uint8_t> bpw, bpf;
        if (!bpd.empty()) {
            bpw.resize(h->ncell);
            bpf.resize(h->nface);
            for (int g = 0; g < n; ++g) {
                Block &B = h->blocks[g];
                const Shape &S = h->shapes[B.shape];
                const int kk = h->kk[g], jj = h->jj[g];
                auto at = [&](int i, int j, int k) {  // 0-based interior, ghosts -1/N
                    return bpd[h->ip3[g] + (k + 2) + (int64_t)(j + 2) * kk +
                               (int64_t)(i + 2) * kk * jj];
                };
                bool all1 = true;
                for (int i = -1; i <= B.ni; ++i)
                    for (int j = -1; j <= B.nj; ++j)
                        for (int k = -1; k <= B.nk; ++k) {
                            const int out = (i < 0 || i >= B.ni) + (j < 0 || j >= B.nj) + (k < 0 || k >= B.nk);
                            if (out > 1) continue;
                            const double v = at(i, j, k);
                            if (v != 0.0 && v != 1.0)
                                throw HipError{"BP must be 0 or 1"};
                            if (v != 1.0) all1 = false;
                        }
                if (all1) B.flags |= 1;
                for (int i = 0; i < B.ni; ++i)
                    for (int j = 0; j < B.nj; ++j)
                        for (int k = 0; k < B.nk; ++k)
                            bpw[B.wo + slot[S.slot_off + k + B.nk * (j + B.nj * i)]] = (uint8_t)at(i, j, k);
                uint8_t *f = &bpf[B.fc];
                for (int j = 0; j < B.nj; ++j)
                    for (int k = 0; k < B.nk; ++k) {
                        f[face_base(0, B.ni, B.nj, B.nk) + fidx_x(k, j, B.nj, B.nk)] = (uint8_t)at(-1, j, k);
                        f[face_base(1, B.ni, B.nj, B.nk) + fidx_x(k, j, B.nj, B.nk)] = (uint8_t)at(B.ni, j, k);
                    }
                for (int i = 0; i < B.ni; ++i)
                    for (int k = 0; k < B.nk; ++k) {
                        f[face_base(2, B.ni, B.nj, B.nk) + fidx_y(k, i, B.ni, B.nk)] = (uint8_t)at(i, -1, k);
                        f[face_base(3, B.ni, B.nj, B.nk) + fidx_y(k, i, B.ni, B.nk)] = (uint8_t)at(i, B.nj, k);
                    }
                for (int i = 0; i < B.ni; ++i)
                    for (int j = 0; j < B.nj; ++j) {
                        f[face_base(4, B.ni, B.nj, B.nk) + fidx_z(j, i, B.ni, B.nj)] = (uint8_t)at(i, j, -1);
                        f[face_base(5, B.ni, B.nj, B.nk) + fidx_z(j, i, B.ni, B.nj)] = (uint8_t)at(i, j, B.nk);
                    }
                if (!all1) h->masked = true;
            }
        } else {
            for (auto &B : h->blocks) B.flags |= 1;
        }

        {
            int dev = 0, lim = 0;
            HC(hipGetDevice(&dev));
            HC(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
            const size_t need = plane_lds_need(*h);
            if (need > (size_t)lim)
                return fail(MGP_ERR_TOPOLOGY, "blocks of %dx%d faces need %zu bytes of LDS per "
                            "workgroup in the plane and layout kernels; the device allows %d",
                            h->max_ni, h->max_nj, need, lim);
        }
        upload(h->d_blocks, h->blocks);
        upload(h->d_shapes, h->shapes);
        upload(h->d_cells, cells);
        upload(h->d_pstart, pstart);
        upload(h->d_slot, slot);
        upload(h->d_items, items);
        upload(h->d_ents, ents);
        upload(h->d_cellx, cellx);
        upload(h->d_fents, fents);
        h->d_zero.alloc(kZeroBytes);
        HC(hipMemsetAsync(h->d_zero.p, 0, kZeroBytes, h->stream));
        std::vector<int> all(n);
        for (int g = 0; g < n; ++g) all[g] = g;
        upload(h->d_blist_all, all);
        for (auto &kv : h->level_blocks) upload(h->d_blist_lvl[kv.first], kv.second);
        if (h->masked) {
            upload(h->d_bpw, bpw);
            upload(h->d_bpf, bpf);
        }
        // BP in the reference layout: the ghostcell operators and the PAR
        // faces of bound_flow read it
        if (!bpd.empty()) upload(h->d_bpsp, bpd);

        // velocity boundary buffers (field buffers at ipbb, src/core/
        // pointers_mod.F90:64-92): kk*jj, kk*ii, jj*ii values per FIX, OP1
        // or PAR face, first index fastest, in grid and face order
        h->bbo.assign(6 * (size_t)n, -1);
        h->nbb = 0;
        for (int g = 0; g < n; ++g)
            for (int f = 0; f < 6; ++f) {
                const int c = h->bc[6 * g + f];
                if (is_flow_bc(c)) h->level_flow.insert(h->level[g]);
                if (c != FIX_BC && c != OP1_BC && c != PAR_BC) continue;
                h->bbo[6 * g + f] = h->nbb;
                h->nbb += f < 2 ? (int64_t)h->kk[g] * h->jj[g]
                        : f < 4 ? (int64_t)h->kk[g] * h->ii[g] : (int64_t)h->jj[g] * h->ii[g];
            }
        upload(h->d_bbo, h->bbo);
        for (DBuf *d : {&h->d_ubnd, &h->d_vbnd, &h->d_wbnd}) {
            d->alloc(std::max<int64_t>(h->nbb, 1) * sizeof(double));
            HC(hipMemsetAsync(d->p, 0, d->n, h->stream));
        }
        h->d_pinf.alloc(6 * (size_t)n * sizeof(double));
        HC(hipMemsetAsync(h->d_pinf.p, 0, h->d_pinf.n, h->stream));

        // 1-D metrics and coefficients (src/ib/noib_mod.F90:141-152,
        // fold-in zeroing :178-244; reciprocals src/core/corefields_mod.F90:
        // 154-173 with divide0)
        auto recip = [](const double *d, int64_t m) {
            std::vector<double> r(m);
            for (int64_t q = 0; q < m; ++q) r[q] = d[q] == 0.0 ? 0.0 : 1.0 / d[q];
            return r;
        };
        upload_raw(h->d_dx, dx, h->n1x * 8);
        upload_raw(h->d_ddx, ddx, h->n1x * 8);
        upload_raw(h->d_ddy, ddy, h->n1y * 8);
        upload_raw(h->d_ddz, ddz, h->n1z * 8);
        upload_raw(h->d_dy, dy, h->n1y * 8);
        upload_raw(h->d_dz, dz, h->n1z * 8);
        upload(h->d_rdx, recip(dx, h->n1x));
        upload(h->d_rdy, recip(dy, h->n1y));
        upload(h->d_rdz, recip(dz, h->n1z));
        upload(h->d_rddx, recip(ddx, h->n1x));
        upload(h->d_rddy, recip(ddy, h->n1y));
        upload(h->d_rddz, recip(ddz, h->n1z));
        {
            const double *d[3] = {dx, dy, dz};
            const int64_t m[3] = {h->n1x, h->n1y, h->n1z};
            for (int a = 0; a < 3; ++a) {
                std::vector<double> r(m[a]);
                for (int64_t q = 0; q < m[a]; ++q) r[q] = 1.0 / (h->cfg.rho * d[a][q]);
                upload(h->d_rpd[a], r);
            }
        }
        std::vector<double> raw[6], fold[6];
        raw[0].assign(h->n1x, 0.0); raw[1].assign(h->n1x, 0.0);
        raw[2].assign(h->n1y, 0.0); raw[3].assign(h->n1y, 0.0);
        raw[4].assign(h->n1z, 0.0); raw[5].assign(h->n1z, 0.0);
        for (int g = 0; g < n; ++g) {
            const double *d[3] = {dx + h->ip1x[g] - 1, dy + h->ip1y[g] - 1, dz + h->ip1z[g] - 1};
            double *lo[3] = {raw[0].data() + h->ip1x[g] - 1, raw[2].data() + h->ip1y[g] - 1, raw[4].data() + h->ip1z[g] - 1};
            double *hi[3] = {raw[1].data() + h->ip1x[g] - 1, raw[3].data() + h->ip1y[g] - 1, raw[5].data() + h->ip1z[g] - 1};
            const int nn[3] = {h->ii[g], h->jj[g], h->kk[g]};
            for (int a = 0; a < 3; ++a)
                for (int q = 3; q <= nn[a] - 2; ++q) {
                    hi[a][q] = 2.0 / ((d[a][q - 1] + d[a][q]) * d[a][q]);
                    lo[a][q] = 2.0 / ((d[a][q - 1] + d[a][q]) * d[a][q - 1]);
                }
        }
        // separable part of GSAP per direction (src/ib/noib_mod.F90:154-157)
        // and the wall fold-in terms per block (:178-244)
        {
            std::vector<double> apd[3];
            const double *dd[3] = {dx, dy, dz};
            const int64_t n1[3] = {h->n1x, h->n1y, h->n1z};
            for (int a = 0; a < 3; ++a) apd[a].assign(n1[a], 0.0);
            for (int g = 0; g < n; ++g) {
                const int64_t o[3] = {h->ip1x[g], h->ip1y[g], h->ip1z[g]};
                const int nn[3] = {h->ii[g], h->jj[g], h->kk[g]};
                for (int a = 0; a < 3; ++a)
                    for (int q = 3; q <= nn[a] - 2; ++q)
                        apd[a][o[a] - 1 + q] = 2.0 / (dd[a][o[a] - 2 + q] * dd[a][o[a] - 1 + q]);
                Block &B = h->blocks[g];
                const int *bc = &h->bc[6 * g];
                B.wallmask = is_wall_lo(bc[0]) | is_wall_hi(bc[1]) << 1 | is_wall_lo(bc[2]) << 2 |
                             is_wall_hi(bc[3]) << 3 | is_wall_lo(bc[4]) << 4 | is_wall_hi(bc[5]) << 5;
                B.wall[0] = raw[0][o[0] + 2];
                B.wall[1] = raw[1][o[0] + 2 + B.ni - 1];
                B.wall[2] = raw[2][o[1] + 2];
                B.wall[3] = raw[3][o[1] + 2 + B.nj - 1];
                B.wall[4] = raw[4][o[2] + 2];
                B.wall[5] = raw[5][o[2] + 2 + B.nk - 1];
            }
            for (int a = 0; a < 3; ++a) upload(h->d_apd[a], apd[a]);
            upload(h->d_blocks, h->blocks);
        }
        for (int q = 0; q < 6; ++q) fold[q] = raw[q];
        for (int g = 0; g < n; ++g) {
            const int *bc = &h->bc[6 * g];
            if (is_wall_lo(bc[0])) fold[0][h->ip1x[g] + 2] = 0.0;
            if (is_wall_hi(bc[1])) fold[1][h->ip1x[g] + h->ii[g] - 3] = 0.0;
            if (is_wall_lo(bc[2])) fold[2][h->ip1y[g] + 2] = 0.0;
            if (is_wall_hi(bc[3])) fold[3][h->ip1y[g] + h->jj[g] - 3] = 0.0;
            if (is_wall_lo(bc[4])) fold[4][h->ip1z[g] + 2] = 0.0;
            if (is_wall_hi(bc[5])) fold[5][h->ip1z[g] + h->kk[g] - 3] = 0.0;
        }
        for (int q = 0; q < 6; ++q) h->fold_host[q] = fold[q];
        for (int q = 0; q < 6; ++q) {
            upload(h->d_raw[q], raw[q]);
            upload(h->d_c64[q], fold[q]);
            if (h->fp32()) {
                std::vector<float> fq(fold[q].begin(), fold[q].end());
                upload(h->d_cT[q], fq);
            }
        }

        // per-cell arrays
        const size_t tsz = h->tsz();
        h->d_ap64.alloc(h->ncell * 8);
        for (int q = 0; q < 7; ++q) h->d_f[q].alloc(h->ncell * tsz);
        h->d_hilf.alloc(h->ncell * tsz);
        h->d_res.alloc(h->ncell * tsz);
        h->d_rhs64.alloc(h->ncell * 8);
        h->d_dp64.alloc(h->ncell * 8);
        if (h->fp32()) {
            h->d_apT.alloc(h->ncell * 4);
            h->d_rhsT.alloc(h->ncell * 4);
        }
        h->d_hface.alloc(h->nface * tsz);
        h->d_rhi[0].alloc(h->nface * tsz);
        h->d_rhi[1].alloc(h->nface * tsz);
        h->d_rlo.alloc(h->nface * tsz);
        if (h->cfg.ityp == 1) {
            h->d_rapT.alloc(h->ncell * tsz);
            h->d_hface2.alloc(h->nface * tsz);
        }
        std::vector<int> parlist;
        for (int b = 0; b < (int)h->blocks.size(); ++b)
            if (h->blocks[b].par) {
                h->level_par.insert(h->blocks[b].level);
                parlist.push_back(b);
            }
        h->npar = (int)parlist.size();
        if (h->npar) upload(h->d_parlist, parlist);
        build_par_ftoc(*h);
        if (!h->level_par.empty()) {
            h->d_hpar.alloc(h->nface * tsz);
            h->d_pbuf.alloc(h->nface * tsz);
            h->d_pbuf64.alloc(h->nface * 8);
        }
        if (h->level_blocks.size() > 1) {
            h->d_hsh.alloc(h->nshell * tsz);
            h->d_dsh.alloc(h->nshell * 8);
        }
        h->d_scal.alloc(64);
        h->d_never.alloc(64);
        HC(hipMemsetAsync(h->d_never.p, 0, 64, h->stream));
        h->d_hist.alloc(std::max(1, h->cfg.nouter) * 8);

        // GSAP, then the SIP factors
        Topo t = h->topo(-1);
        launch_ap(t, h->d_dx.as<double>(), h->d_dy.as<double>(),
                  h->d_dz.as<double>(), h->d_raw[0].as<double>(),
                  h->d_raw[1].as<double>(), h->d_raw[2].as<double>(),
                  h->d_raw[3].as<double>(), h->d_raw[4].as<double>(),
                  h->d_raw[5].as<double>(), h->mask(), h->d_ap64.as<double>(),
                  h->stream);
        HC(hipGetLastError());
        if (h->fp32()) {
            setup_factors<float>(*h);
            share_factors<float>(*h);
            launch_convert<float>(h->d_ap64.as<double>(), h->d_apT.as<float>(), h->ncell, h->stream);
        } else {
            setup_factors<double>(*h);
            share_factors<double>(*h);
        }
        if (h->cfg.ityp == 1) {
            if (h->fp32()) launch_rap<float>(h->d_ap64.as<double>(), h->d_rapT.as<float>(), h->ncell, h->stream);
            else launch_rap<double>(h->d_ap64.as<double>(), h->d_rapT.as<double>(), h->ncell, h->stream);
            HC(hipGetLastError());
        }
        HC(hipStreamSynchronize(h->stream));
        const char *ps = getenv("MGP_PLAN_SHAPE");
        const int forced = ps ? atoi(ps) : 0;
        if (forced >= 1 && forced <= kNumShapes) {
            h->plan_shape = h->plan_shape_bwd = forced;
        } else {
            if (h->fp32()) tune_shape<float>(*h);
            else tune_shape<double>(*h);
        }
    } catch (const HipError &e) {
        return fail(MGP_ERR_HIP, "%s", e.msg.c_str());
    }
    *out = h.release();
    g_err.clear();
    return MGP_OK;
}

// Autotuner of the plane kernels' pipeline shape: times, per shape, the
// kernels of one outer iteration (ninner sweeps or relaxations and one
// outer update) on the handle's blocks and keeps the fastest; the backward
// sweep, timed on its own, gets the shape fastest for it.  shape_ms holds
// the outer iteration's time with every kernel in one shape.  Runs on
// whatever the work arrays hold (before the first solve: scratch); the
// solver state is reset at the start of every solve.
template <typename T>
void tune_shape(mgp_handle &h) {
    const hipStream_t s = h.stream;
    Topo tall = h.topo(-1);
    Fields<T> F = h.fields<T>();
    Mask m = h.mask();
    Outer o = h.outer();
    o.done = (int *)h.d_never.p;
    const int *never = (const int *)h.d_never.p;
    const bool classic = h.cfg.ityp == 2;
    const bool sor = h.cfg.ityp == 1;
    T *rhsT = h.fp32() ? h.d_rhsT.as<T>() : nullptr;
    hipEvent_t e0, e1;
    HC(hipEventCreate(&e0));
    HC(hipEventCreate(&e1));
    auto fwd = [&]() {
        if (sor) launch_sor<T>(tall, F, m, F.hface, h.d_hface2.as<T>(), h.cfg.omega, 0, never, s);
        else launch_forward<T>(tall, F, m, 0, 1, 0, classic, never, s);
    };
    auto bwd = [&]() {
        if (!sor) launch_backward<T>(tall, F, 0, classic, never, s);
    };
    auto timed = [&](auto &&f, int reps) {
        float ms = 0.f;
        HC(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) f();
        HC(hipEventRecord(e1, s));
        HC(hipEventSynchronize(e1));
        HC(hipEventElapsedTime(&ms, e0, e1));
        return (double)ms;
    };
    const int nin = std::max(1, h.cfg.ninner);
    int best = 0, best_b = 0;
    double tf[kNumShapes], tb[kNumShapes];
    for (int v = 1; v <= kNumShapes; ++v) {
        set_plan_shape(v);
        fwd();  // warm-up
        bwd();
        const double a = timed(fwd, 2), ab = timed(bwd, 2);
        const double b = timed([&]() { launch_outer<T>(tall, F, m, o, rhsT, s); }, 1);
        tf[v - 1] = 0.5 * a * nin + b;
        tb[v - 1] = 0.5 * ab * nin;
        h.shape_ms[v - 1] = tf[v - 1] + tb[v - 1];
        if (!best || tf[v - 1] < tf[best - 1]) best = v;
        if (!best_b || tb[v - 1] < tb[best_b - 1]) best_b = v;
    }
    HC(hipGetLastError());
    HC(hipEventDestroy(e0));
    HC(hipEventDestroy(e1));
    h.plan_shape = best;
    h.plan_shape_bwd = sor ? best : best_b;
    set_plan_shapes(h.plan_shape, h.plan_shape_bwd);
}

// Pool fields start at zero (src/core/fieldpool_mod.F90:113-122).
template <typename T>
void reset_state(mgp_handle &h) {
    const hipStream_t s = h.stream;
    launch_fill<double>(h.d_dp64.as<double>(), h.ncell, 0.0, s);
    launch_fill<T>(h.d_hilf.as<T>(), h.ncell, T(0), s);
    launch_fill<T>(h.d_hface.as<T>(), h.nface, T(0), s);
    launch_fill<T>(h.d_rhi[0].as<T>(), h.nface, T(0), s);
    launch_fill<T>(h.d_rhi[1].as<T>(), h.nface, T(0), s);
    launch_fill<T>(h.d_rlo.as<T>(), h.nface, T(0), s);
    if (h.d_hface2.p) launch_fill<T>(h.d_hface2.as<T>(), h.nface, T(0), s);
    if (h.d_hpar.p) launch_fill<T>(h.d_hpar.as<T>(), h.nface, T(0), s);
    if (h.d_hsh.p) {
        launch_fill<T>(h.d_hsh.as<T>(), h.nshell, T(0), s);
        launch_fill<double>(h.d_dsh.as<double>(), h.nshell, 0.0, s);
    }
}

// The outer defect-correction loop (src/flow/pressuresolver_mod.F90:
// 253-303) on the RHS already in place, with the operator set (F, o):
// the forward operators, or the transposed ones for the adjoint.  Returns
// the device time of the loop in ms.
template <typename T>
float iterate(mgp_handle &h, const Fields<T> &F, const Outer &o,
              double prefak) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "T");
    const mgp_config &c = h.cfg;
    const hipStream_t s = h.stream;
    set_plan_shapes(h.plan_shape, h.plan_shape_bwd);
    Topo tall = h.topo(-1);
    T *rhsT = h.fp32() ? h.d_rhsT.as<T>() : nullptr;
    hipEvent_t e0, e1;
    HC(hipEventCreate(&e0));
    HC(hipEventCreate(&e1));
    HC(hipEventRecord(e0, s));
    if (c.nouter > 0) {
        // chunk of outer iterations that returns the RES buffer parity
        int chunk = std::max(1, c.check_every);
        if (((c.ninner - 1) & 1) && (chunk & 1)) ++chunk;
        int launched = 0;
        const bool classic = c.ityp == 2;
        const bool multi = h.level_blocks.size() > 1;
        const int coarsest = h.level_blocks.begin()->first;
        Outer o_inner = o;
        if (multi) o_inner.gmax = (unsigned long long *)(h.scal() + 5);  // unused slot
        const double *mt[6] = {h.d_dx.as<double>(), h.d_dy.as<double>(), h.d_dz.as<double>(),
                               h.d_ddx.as<double>(), h.d_ddy.as<double>(), h.d_ddz.as<double>()};
        const double *dd[3] = {mt[3], mt[4], mt[5]};
        T *pbuf = h.d_pbuf.as<T>();
        const int bpv = c.ibtype == 1;
        // one level of SIP sweeps: the outer update runs fused with the next
        // outer iteration's first forward sweep (MGP_FUSED_OUTER=0: apart)
        const char *fe = getenv("MGP_FUSED_OUTER");
        const bool fuse = !multi && !(c.ityp == 1 && c.sor_all_levels) && !(fe && atoi(fe) == 0);
        auto enqueue_chunk = [&]() {
            // RES buffer parity per level (each level exchanges its own RES)
            std::map<int, int> rd;
            Mask m = h.mask();
            for (int it = 0; it < chunk; ++it) {
                // coarse to fine: ctof, parent, mgpoisit
                // (src/flow/pressuresolver_mod.F90:224-231, 361-420)
                for (auto &kv : h.level_blocks) {
                    const int L = kv.first;
                    Topo t = h.topo(L);
                    const bool first = L == coarsest;
                    const bool par = h.level_par.count(L) > 0;
                    const bool sor = c.ityp == 1 && (L > h.minlevel || c.sor_all_levels);
                    T *hf[2] = {F.hface, h.d_hface2.as<T>()};
                    if (!first) {
                        // the first relaxation or sweep reads its ghosts from
                        // the buffer ctof fills
                        launch_ctof<T>(t, F, sor ? hf[c.ninner & 1] : F.hface,
                                       h.d_hsh.as<T>(), h.d_dsh.as<double>(), o.done, s);
                        if (par) launch_parent<T>(t, F, pbuf, o.done, s);
                    }
                    if (sor) {
                        // the last relaxation writes the faces into F.hface,
                        // where the SIP levels leave theirs too
                        for (int il = 1; il <= c.ninner; ++il) {
                            if (par) launch_bound<T>(t, F, pbuf, m, bpv, mt, o.done, s);
                            launch_sor<T>(t, F, m, hf[(c.ninner - il + 1) & 1],
                                          hf[(c.ninner - il) & 1], c.omega,
                                          il == 1 && first, o.done, s);
                        }
                    } else {
                        int &r = rd[L];
                        for (int il = 1; il <= c.ninner; ++il) {
                            const int last = il == c.ninner;
                            if (par) launch_bound<T>(t, F, pbuf, m, bpv, mt, o.done, s);
                            // fused: the previous outer update ran this sweep
                            if (!(fuse && il == 1))
                                launch_forward<T>(t, F, m, r, !last, il == 1 && first, classic,
                                                  o.done, s);
                            launch_backward<T>(t, F, il == 1 && first, classic, o.done, s);
                            if (!last) r ^= 1;
                        }
                    }
                    if (par) launch_bound<T>(t, F, pbuf, m, bpv, mt, o.done, s);
                }
                // fine to coarse: HILF into the parents, the residual, the
                // RHS into the parents, max |RHS| (:233-298)
                for (auto lv = h.level_blocks.rbegin(); lv != h.level_blocks.rend(); ++lv)
                    if (lv->first != coarsest)
                        launch_restrict<T, T>(h.topo(lv->first), F.hilf, F.hilf, F.hface,
                                              nullptr, m, bpv ? 1 : 0, dd, o.done, s);
                if (fuse)   // with the next iteration's first sweep (RES parity rd)
                    launch_outer_fwd<T>(tall, F, m, o_inner, rhsT, rd[coarsest], c.ninner > 1,
                                        classic, s);
                else
                    launch_outer<T>(tall, F, m, o_inner, rhsT, s);
                if (multi) {
                    for (auto lv = h.level_blocks.rbegin(); lv != h.level_blocks.rend(); ++lv)
                        if (lv->first != coarsest)
                            launch_restrict<double, double>(h.topo(lv->first), o.rhs64, o.rhs64,
                                                            nullptr, h.fp32() ? h.d_rhsT.as<float>()
                                                                              : nullptr,
                                                            m, bpv ? 2 : 0, dd, o.done, s);
                    launch_maxabs_acc(tall, o.rhs64, o, s);
                }
                launch_check(o, c.nouter, c.nouter_min, c.epcorr, c.eps_rel,
                             prefak, h.d_hist.as<double>(), c.nouter, s);
            }
        };
        // capture one chunk into a graph; streams that cannot be captured
        // (the legacy null stream) get the same launches directly
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        if (s != nullptr && hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
            enqueue_chunk();
            HC(hipStreamEndCapture(s, &graph));
            HC(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        } else {
            (void)hipGetLastError();
        }
        // optional progress line on stderr (MGP_PROGRESS_SECONDS > 0)
        const char *pe = getenv("MGP_PROGRESS_SECONDS");
        const double every = pe ? atof(pe) : 0.0;
        auto t_last = std::chrono::steady_clock::now();
        int hstate[2] = {0, 0};   // outer iterations done, done flag
        // fused: the first forward sweep of the first outer iteration
        if (fuse)
            launch_forward<T>(h.topo(coarsest), F, h.mask(), 0, c.ninner > 1, 1, classic,
                              o.done, s);
        while (launched < c.nouter) {
            if (exec) HC(hipGraphLaunch(exec, s));
            else enqueue_chunk();
            launched += chunk;
            if (launched >= c.nouter) break;
            HC(hipMemcpyAsync(hstate, o.ctr, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
            HC(hipStreamSynchronize(s));
            if (every > 0.0) {
                const auto now = std::chrono::steady_clock::now();
                if (std::chrono::duration<double>(now - t_last).count() >= every) {
                    double mr[2];
                    HC(hipMemcpy(mr, o.maxrhs0, 2 * sizeof(double), hipMemcpyDeviceToHost));
                    fprintf(stderr, "[mgp] outer %d (inner %d): max|RHS| %.3e, relative %.3e\n",
                            hstate[0], hstate[0] * c.ninner, mr[1], mr[1] / mr[0]);
                    fflush(stderr);
                    t_last = now;
                }
            }
            if (hstate[1]) break;
        }
        if (exec) HC(hipGraphExecDestroy(exec));
        if (graph) HC(hipGraphDestroy(graph));
    }
    HC(hipEventRecord(e1, s));
    HC(hipEventSynchronize(e1));
    float ms = 0.f;
    HC(hipEventElapsedTime(&ms, e0, e1));
    HC(hipEventDestroy(e0));
    HC(hipEventDestroy(e1));
    return ms;
}

// DP (or the adjoint's y) to the reference layout in h.d_dpsp, with the
// face ghosts of a 1-layer exchange (src/flow/pressuresolver_mod.F90:278
// accumulates HILF including the ghosts that conn(...) copied).
void dp_to_reference(mgp_handle &h) {
    const hipStream_t s = h.stream;
    Topo tall = h.topo(-1);
    if (h.d_dpsp.n < (size_t)h.n3d * 8) h.d_dpsp.alloc((size_t)h.n3d * 8);
    launch_wo_to_sp<double>(tall, h.d_dp64.as<double>(), h.d_dpsp.as<double>(), s, true);
    if (h.d_dsh.p) launch_shell_to_sp(tall, h.d_dsh.as<double>(), h.d_dpsp.as<double>(), s);
    double *fd[1] = {h.d_dpsp.as<double>()};
    exchange(h, -1, 1, 0, 0, fd, 1);
}

int finish_stats(mgp_handle &h, float ms, mgp_stats *st) {
    const mgp_config &c = h.cfg;
    double sc[3];
    int iv[4];
    HC(hipMemcpyAsync(sc, h.d_scal.p, 3 * sizeof(double), hipMemcpyDeviceToHost, h.stream));
    HC(hipMemcpyAsync(iv, (char *)h.d_scal.p + 24, 4 * sizeof(int), hipMemcpyDeviceToHost, h.stream));
    HC(hipStreamSynchronize(h.stream));
    h.have_solution = true;
    if (st) {
        st->nouter_done = iv[0];
        st->ipc = iv[0] * c.ninner;
        st->converged = iv[2];
        st->maxrhs0 = sc[1];
        st->maxrhs = sc[2];
        st->solve_ms = ms;
    }
    if (iv[3]) return fail(MGP_ERR_NONFINITE, "residual became NaN or Inf");
    return MGP_OK;
}

template <typename T>
int solve(mgp_handle &h, double *u, double *v, double *w, double *p,
          double dt, mgp_stats *st) {
    const mgp_config &c = h.cfg;
    const double prefak = c.rho / dt;
    const hipStream_t s = h.stream;
    Topo tall = h.topo(-1);
    T *rhsT = h.fp32() ? h.d_rhsT.as<T>() : nullptr;
    reset_state<T>(h);
    launch_divcal<T>(tall, u, v, w, h.d_rddx.as<double>(), h.d_rddy.as<double>(),
                     h.d_rddz.as<double>(), c.ibtype == 1 ? h.d_bpsp.as<double>() : nullptr,
                     h.sdiv_sp(), prefak, h.d_rhs64.as<double>(), rhsT, s);
    // RHS of the refined grids into their parents (:197-199)
    const double *dd[3] = {h.d_ddx.as<double>(), h.d_ddy.as<double>(), h.d_ddz.as<double>()};
    for (auto lv = h.level_blocks.rbegin(); lv != h.level_blocks.rend(); ++lv)
        if (lv->first != h.level_blocks.begin()->first)
            launch_restrict<double, double>(h.topo(lv->first), h.d_rhs64.as<double>(),
                                            h.d_rhs64.as<double>(), nullptr,
                                            h.fp32() ? h.d_rhsT.as<float>() : nullptr, h.mask(),
                                            c.ibtype == 1 ? 2 : 0, dd, h.d_never.as<int>(), s);
    launch_maxabs(tall, h.d_rhs64.as<double>(), h.outer(), s);
    HC(hipGetLastError());
    const float ms = iterate<T>(h, h.fields<T>(), h.outer(), prefak);

    // DP -> reference layout; parent and bound_pressure of DP per level,
    // mgpcorr, U, V, W, P into the parents and the post-projection
    // exchange (src/flow/pressuresolver_mod.F90:323-348)
    dp_to_reference(h);
    const double *mt[6] = {h.d_dx.as<double>(), h.d_dy.as<double>(), h.d_dz.as<double>(),
                           h.d_ddx.as<double>(), h.d_ddy.as<double>(), h.d_ddz.as<double>()};
    for (int L : h.level_par) {
        launch_parent_sp(h.topo(L), h.d_dpsp.as<double>(), h.d_pbuf64.as<double>(), s);
        launch_bound_sp(h.topo(L), h.d_dpsp.as<double>(), h.d_pbuf64.as<double>(),
                        c.ibtype == 1 ? h.d_bpsp.as<double>() : nullptr, c.ibtype == 1, mt, s);
    }
    // on single-level grids the correction can write the k-face ghosts of
    // the closing exchange itself (kpush_on)
    const bool kpush = !c.sync_by_caller && kpush_on(h);
    const CorrChains cc{h.d_cchain.as<int>(), h.d_cchains.as<int2>(), h.ncchain, h.ccnx};
    launch_mgpcorr(tall, u, v, w, p, h.d_dpsp.as<double>(), h.d_rdx.as<double>(),
                   h.d_rdy.as<double>(), h.d_rdz.as<double>(), dt / c.rho,
                   c.ibtype == 1 ? h.d_bpsp.as<double>() : nullptr, kpush ? &cc : nullptr, s);
    for (auto lv = h.level_blocks.rbegin(); lv != h.level_blocks.rend(); ++lv)
        if (lv->first != h.level_blocks.begin()->first)
            launch_ftoc_uvwp(h.topo(lv->first), u, v, w, p, c.ibtype == 1,
                             c.ibtype == 1 ? h.d_bpsp.as<double>() : nullptr, dd, s);
    // per level, coarse to fine: U, V, W at the PAR faces from the parent,
    // bound_flow, and the exchange inside the level, which takes the ghost
    // values bound_flow set (:338-348); a caller that runs these itself
    // asks to stop after ftoc
    if (!c.sync_by_caller) {
        double *fs[4] = {u, v, w, p};
        for (int l = h.minlevel; l <= h.maxlevel; ++l) {
            if (!h.level_blocks.count(l)) continue;
            const Topo tl = h.topo(l);
            if (h.level_par.count(l))
                launch_parent_uvw(tl, u, v, w, h.d_ubnd.as<double>(), h.d_vbnd.as<double>(),
                                  h.d_wbnd.as<double>(), h.d_bbo.as<int64_t>(), s);
            if (h.level_flow.count(l))
                launch_bound_flow(tl, u, v, w, p, h.d_ubnd.as<double>(), h.d_vbnd.as<double>(),
                                  h.d_wbnd.as<double>(), h.d_bbo.as<int64_t>(),
                                  h.d_pinf.as<double>(), h.d_bpsp.as<double>(), dd[0], dd[1],
                                  dd[2], c.rho, s);
            exchange(h, l, 2, 1, 0, fs, 4, kpush);
        }
    }
    HC(hipGetLastError());
    return finish_stats(h, ms, st);
}

// Operands of the transposed problem A^T y = b, built on first use: the
// 1-D coefficients of A^T (the coefficient of cell c towards neighbour n
// is A's coefficient of n towards c: awT(i) = GSAE(i-1), taken from the
// neighbouring block at the block's first cell, zero without a connected
// neighbour) and the transposed SIP factors (launch_tfactors).
template <typename T>
void ensure_adjoint(mgp_handle &h) {
    if (h.adj_ready) return;
    const int n = h.ngrid;
    std::vector<double> tr[6];
    for (int q = 0; q < 6; ++q) tr[q].assign(h.fold_host[q].size(), 0.0);
    const int64_t *o1[3] = {h.ip1x.data(), h.ip1y.data(), h.ip1z.data()};
    for (int g = 0; g < n; ++g) {
        const Block &B = h.blocks[g];
        const int nn[3] = {h.ii[g], h.jj[g], h.kk[g]};
        for (int a = 0; a < 3; ++a) {
            const std::vector<double> &lo = h.fold_host[2 * a], &hi = h.fold_host[2 * a + 1];
            double *tlo = tr[2 * a].data() + o1[a][g] - 1;   // 1-based index q
            double *thi = tr[2 * a + 1].data() + o1[a][g] - 1;
            const double *flo = lo.data() + o1[a][g] - 1, *fhi = hi.data() + o1[a][g] - 1;
            const int last = nn[a] - 2;
            for (int q = 3; q <= last; ++q) {
                if (q > 3) tlo[q] = fhi[q - 1];
                else if (B.nbr[2 * a] >= 0) {
                    const int nb = B.nbr[2 * a];
                    const int nlast = (a == 0 ? h.ii[nb] : a == 1 ? h.jj[nb] : h.kk[nb]) - 2;
                    tlo[q] = hi[o1[a][nb] - 1 + nlast];
                }
                if (q < last) thi[q] = flo[q + 1];
                else if (B.nbr[2 * a + 1] >= 0) {
                    const int nb = B.nbr[2 * a + 1];
                    thi[q] = lo[o1[a][nb] - 1 + 3];
                }
            }
        }
    }
    for (int q = 0; q < 6; ++q) {
        upload(h.d_c64t[q], tr[q]);
        if (h.fp32()) {
            std::vector<float> fq(tr[q].begin(), tr[q].end());
            upload(h.d_cTt[q], fq);
        }
    }
    const size_t tsz = h.tsz();
    for (int q = 0; q < 6; ++q) h.d_ft[q].alloc(h.nfcell * tsz);
    const T *f[7];
    T *g[6];
    for (int q = 0; q < 7; ++q) f[q] = h.d_f[q].as<T>();
    for (int q = 0; q < 6; ++q) g[q] = h.d_ft[q].as<T>();
    launch_tfactors<T>(h.topo_reps(), h.mask(), f, g, h.stream);
    HC(hipGetLastError());
    HC(hipStreamSynchronize(h.stream));
    h.adj_ready = true;
}

// Adjoint of the projection (no counterpart in the reference): on entry
// u, v, w hold dJ/du' on the interior faces, on exit dJ/du.
template <typename T>
int solve_adjoint(mgp_handle &h, double *u, double *v, double *w, double dt,
                  mgp_stats *st) {
    ensure_adjoint<T>(h);
    const mgp_config &c = h.cfg;
    const double prefak = c.rho / dt;
    const hipStream_t s = h.stream;
    Topo tall = h.topo(-1);
    T *rhsT = h.fp32() ? h.d_rhsT.as<T>() : nullptr;
    const double *bpsp = c.ibtype == 1 ? h.d_bpsp.as<double>() : nullptr;
    reset_state<T>(h);
    launch_adj_clean(tall, u, v, w, bpsp, s);
    double *fs[3] = {u, v, w};
    exchange(h, -1, 1, 0, 0, fs, 3);
    launch_gradT<T>(tall, u, v, w, h.d_rdx.as<double>(), h.d_rdy.as<double>(),
                    h.d_rdz.as<double>(), bpsp, h.d_rhs64.as<double>(), rhsT, s);
    Outer o = h.outer_t();
    launch_maxabs(tall, h.d_rhs64.as<double>(), o, s);
    HC(hipGetLastError());
    const float ms = iterate<T>(h, h.fields_t<T>(), o, prefak);
    dp_to_reference(h);
    launch_divT(tall, u, v, w, h.d_dpsp.as<double>(), h.d_rddx.as<double>(),
                h.d_rddy.as<double>(), h.d_rddz.as<double>(), bpsp, s);
    launch_adj_clean(tall, u, v, w, bpsp, s);
    HC(hipGetLastError());
    return finish_stats(h, ms, st);
}

int adjoint_any(mgp_handle *h, double *u, double *v, double *w, double dt,
                mgp_stats *st) {
    if (!(dt > 0.0) && !(dt < 0.0)) return fail(MGP_ERR_ARG, "dt must be non-zero");
    if (h->masked && h->cfg.ibtype != 1)
        return fail(MGP_ERR_ARG, "the adjoint of a masked operator needs ibtype 1 "
                    "(ghostcell divcal and mgpcorr)");
    if (h->level_blocks.size() > 1)
        return fail(MGP_ERR_TOPOLOGY, "the adjoint supports one grid level");
    try {
        int rc = h->fp32() ? solve_adjoint<float>(*h, u, v, w, dt, st)
                           : solve_adjoint<double>(*h, u, v, w, dt, st);
        if (rc == MGP_OK) g_err.clear();
        return rc;
    } catch (const HipError &e) {
        return fail(MGP_ERR_HIP, "%s", e.msg.c_str());
    }
}

int solve_any(mgp_handle *h, double *u, double *v, double *w, double *p,
              double dt, mgp_stats *st) {
    if (!(dt > 0.0) && !(dt < 0.0)) return fail(MGP_ERR_ARG, "dt must be non-zero");
    try {
        int rc = h->fp32() ? solve<float>(*h, u, v, w, p, dt, st)
                           : solve<double>(*h, u, v, w, p, dt, st);
        if (rc == MGP_OK) g_err.clear();
        return rc;
    } catch (const HipError &e) {
        return fail(MGP_ERR_HIP, "%s", e.msg.c_str());
    }
}

}  // namespace

// =====================================================================
extern "C" {

int mgp_init(const mgp_grids *grids, const mgp_config *cfg,
             const mgletreal *dx, const mgletreal *dy, const mgletreal *dz,
             const mgletreal *ddx, const mgletreal *ddy, const mgletreal *ddz,
             const mgletreal *bp, mgp_handle **out) {
    MGP_RANGE("mgp_init");
    return init_common(grids, cfg, dx, dy, dz, ddx, ddy, ddz, bp, nullptr, out);
}

int mgp_init_mask(const mgp_grids *grids, const mgp_config *cfg,
                  const mgletreal *dx, const mgletreal *dy,
                  const mgletreal *dz, const mgletreal *ddx,
                  const mgletreal *ddy, const mgletreal *ddz,
                  const int8_t *blocked, mgp_handle **out) {
    MGP_RANGE("mgp_init_mask");
    return init_common(grids, cfg, dx, dy, dz, ddx, ddy, ddz, nullptr,
                       blocked, out);
}

int mgp_mgpoisl_device(mgp_handle *h, mgletreal *u, mgletreal *v,
                       mgletreal *w, mgletreal *p, mgletreal dt,
                       mgletint ittot, mgletint irk, mgp_stats *stats) {
    MGP_RANGE("mgp_mgpoisl_device");
    (void)ittot;
    (void)irk;
    if (!h) return fail(MGP_ERR_STATE, "handle not initialised");
    if (!u || !v || !w || !p) return fail(MGP_ERR_ARG, "null field");
    return solve_any(h, u, v, w, p, dt, stats);
}

int mgp_mgpoisl(mgp_handle *h, mgletreal *u, mgletreal *v, mgletreal *w,
                mgletreal *p, mgletreal dt, mgletint ittot, mgletint irk,
                mgp_stats *stats) {
    MGP_RANGE("mgp_mgpoisl");
    if (!h) return fail(MGP_ERR_STATE, "handle not initialised");
    if (!u || !v || !w || !p) return fail(MGP_ERR_ARG, "null field");
    try {
        const size_t bytes = (size_t)h->n3d * 8;
        DBuf *d[4] = {&h->d_u, &h->d_v, &h->d_w, &h->d_p};
        double *hp[4] = {u, v, w, p};
        for (int q = 0; q < 4; ++q) {
            if (d[q]->n < bytes) d[q]->alloc(bytes);
            HC(hipMemcpyAsync(d[q]->p, hp[q], bytes, hipMemcpyHostToDevice, h->stream));
        }
        int rc = solve_any(h, h->d_u.as<double>(), h->d_v.as<double>(),
                           h->d_w.as<double>(), h->d_p.as<double>(), dt, stats);
        for (int q = 0; q < 4; ++q)
            HC(hipMemcpyAsync(hp[q], d[q]->p, bytes, hipMemcpyDeviceToHost, h->stream));
        HC(hipStreamSynchronize(h->stream));
        return rc;
    } catch (const HipError &e) {
        return fail(MGP_ERR_HIP, "%s", e.msg.c_str());
    }
}

int mgp_adjoint_device(mgp_handle *h, mgletreal *ubar, mgletreal *vbar,
                       mgletreal *wbar, mgletreal dt, mgp_stats *stats) {
    MGP_RANGE("mgp_adjoint_device");
    if (!h) return fail(MGP_ERR_STATE, "handle not initialised");
    if (!ubar || !vbar || !wbar) return fail(MGP_ERR_ARG, "null field");
    return adjoint_any(h, ubar, vbar, wbar, dt, stats);
}

int mgp_adjoint(mgp_handle *h, mgletreal *ubar, mgletreal *vbar,
                mgletreal *wbar, mgletreal dt, mgp_stats *stats) {
    MGP_RANGE("mgp_adjoint");
    if (!h) return fail(MGP_ERR_STATE, "handle not initialised");
    if (!ubar || !vbar || !wbar) return fail(MGP_ERR_ARG, "null field");
    try {
        const size_t bytes = (size_t)h->n3d * 8;
        DBuf *d[3] = {&h->d_u, &h->d_v, &h->d_w};
        double *hp[3] = {ubar, vbar, wbar};
        for (int q = 0; q < 3; ++q) {
            if (d[q]->n < bytes) d[q]->alloc(bytes);
            HC(hipMemcpyAsync(d[q]->p, hp[q], bytes, hipMemcpyHostToDevice, h->stream));
        }
        int rc = adjoint_any(h, h->d_u.as<double>(), h->d_v.as<double>(),
                             h->d_w.as<double>(), dt, stats);
        for (int q = 0; q < 3; ++q)
            HC(hipMemcpyAsync(hp[q], d[q]->p, bytes, hipMemcpyDeviceToHost, h->stream));
        HC(hipStreamSynchronize(h->stream));
        return rc;
    } catch (const HipError &e) {
        return fail(MGP_ERR_HIP, "%s", e.msg.c_str());
    }
}

int mgp_finish(mgp_handle *h) {
    MGP_RANGE("mgp_finish");
    if (!h) return fail(MGP_ERR_STATE, "handle not initialised");
    (void)hipStreamSynchronize(h->stream);
    delete h;
    g_err.clear();
    return MGP_OK;
}

const char *mgp_last_error(void) { return g_err.c_str(); }

int mgp_set_iterations(mgp_handle *h, mgletint ninner, mgletint nouter,
                       mgletint nouter_min, mgletreal epcorr,
                       mgletreal eps_rel, mgletint check_every) {
    MGP_RANGE("mgp_set_iterations");
    if (!h) return fail(MGP_ERR_STATE, "handle not initialised");
    mgp_config c = h->cfg;
    c.ninner = ninner;
    c.nouter = nouter;
    c.nouter_min = nouter_min;
    c.epcorr = epcorr;
    c.eps_rel = eps_rel;
    c.check_every = check_every < 1 ? 1 : check_every;
    int rc = check_cfg(c);
    if (rc) return rc;
    h->cfg = c;
    try {
        if (h->d_hist.n < (size_t)std::max(1, nouter) * 8)
            h->d_hist.alloc((size_t)std::max(1, nouter) * 8);
    } catch (const HipError &e) {
        return fail(MGP_ERR_HIP, "%s", e.msg.c_str());
    }
    return MGP_OK;
}

int mgp_set_stream(mgp_handle *h, void *stream) {
    MGP_RANGE("mgp_set_stream");
    if (!h) return fail(MGP_ERR_STATE, "handle not initialised");
    h->stream = stream ? (hipStream_t)stream : h->own_stream;
    return MGP_OK;
}

int mgp_get_field(mgp_handle *h, const char *name, mgletreal *out) {
    MGP_RANGE("mgp_get_field");
    return mgp_get_field_grid(h, name, 0, out);
}

int mgp_get_field_grid(mgp_handle *h, const char *name, mgletint igrid, mgletreal *out) {
    MGP_RANGE("mgp_get_field_grid");
    if (!h || !name || !out) return fail(MGP_ERR_ARG, "null argument");
    if (igrid < 0 || igrid > h->ngrid) return fail(MGP_ERR_ARG, "invalid grid %d", igrid);
    const std::string nm(name);
    try {
        const hipStream_t s = h->stream;
        const char *c1[6] = {"GSAW", "GSAE", "GSAS", "GSAN", "GSAB", "GSAT"};
        for (int q = 0; q < 6; ++q)
            if (nm == c1[q]) {
                if (igrid > 0) return fail(MGP_ERR_ARG, "1-D fields are copied whole");
                HC(hipMemcpyAsync(out, h->d_c64[q].p, h->d_c64[q].n, hipMemcpyDeviceToHost, s));
                HC(hipStreamSynchronize(s));
                return MGP_OK;
            }
        DBuf tmp;
        tmp.alloc((size_t)h->n3d * 8);
        launch_fill<double>(tmp.as<double>(), h->n3d, 0.0, s);
        Topo t = h->topo(-1);
        const char *fn[7] = {"SIPLW", "SIPLS", "SIPLB", "SIPLPR", "SIPUE", "SIPUN", "SIPUT"};
        bool found = false;
        for (int q = 0; q < 7; ++q)
            if (nm == fn[q]) {
                if (h->fp32())
                    launch_to_ref_factors<float>(t, h->d_f[q].as<float>(), tmp.as<double>(), h->cfg.ityp == 2, s);
                else
                    launch_to_ref_factors<double>(t, h->d_f[q].as<double>(), tmp.as<double>(), h->cfg.ityp == 2, s);
                found = true;
            }
        if (nm == "GSAP") {
            launch_wo_to_sp<double>(t, h->d_ap64.as<double>(), tmp.as<double>(), s);
            found = true;
        } else if (nm == "DPG") {
            if (!h->have_solution) return fail(MGP_ERR_STATE, "no mgpoisl call yet");
            HC(hipMemcpyAsync(tmp.p, h->d_dpsp.p, (size_t)h->n3d * 8, hipMemcpyDeviceToDevice, s));
            found = true;
        } else if (nm == "DP" || nm == "RHS") {
            if (!h->have_solution) return fail(MGP_ERR_STATE, "no mgpoisl call yet");
            launch_wo_to_sp<double>(t, nm == "DP" ? h->d_dp64.as<double>() : h->d_rhs64.as<double>(), tmp.as<double>(), s);
            found = true;
        } else if (nm == "BP") {
            if (h->masked) {
                launch_fill<double>(tmp.as<double>(), h->n3d, 1.0, s);
                launch_wo_to_sp<uint8_t>(t, h->d_bpw.as<uint8_t>(), tmp.as<double>(), s);
            } else {
                launch_fill<double>(tmp.as<double>(), h->n3d, 1.0, s);
            }
            found = true;
        }
        if (!found) return fail(MGP_ERR_ARG, "unknown field %s", name);
        if (igrid > 0) {
            const int b = igrid - 1;
            const size_t n = (size_t)h->kk[b] * h->jj[b] * h->ii[b];
            HC(hipMemcpyAsync(out, tmp.as<double>() + h->ip3[b], n * 8, hipMemcpyDeviceToHost, s));
        } else {
            HC(hipMemcpyAsync(out, tmp.p, (size_t)h->n3d * 8, hipMemcpyDeviceToHost, s));
        }
        HC(hipStreamSynchronize(s));
    } catch (const HipError &e) {
        return fail(MGP_ERR_HIP, "%s", e.msg.c_str());
    }
    return MGP_OK;
}

int mgp_connect_device(mgp_handle *h, mgletint level, mgletint layers,
                       mgletint corners, mgletint forward,
                       mgletreal **fields, mgletint nfield) {
    MGP_RANGE("mgp_connect_device");
    if (!h) return fail(MGP_ERR_STATE, "handle not initialised");
    if (layers < 1 || layers > 2) return fail(MGP_ERR_ARG, "layers must be 1 or 2");
    if (forward != 0 && corners) return fail(MGP_ERR_ARG, "forward with corners");
    if (forward < -1 || forward > 1) return fail(MGP_ERR_ARG, "forward must be -1, 0 or 1");
    if (!fields || nfield < 1) return fail(MGP_ERR_ARG, "no fields");
    try {
        exchange(*h, level, layers, corners, forward, fields, nfield);
        HC(hipGetLastError());
        HC(hipStreamSynchronize(h->stream));
    } catch (const HipError &e) {
        return fail(MGP_ERR_HIP, "%s", e.msg.c_str());
    }
    return MGP_OK;
}

int mgp_par_ftoc_device(mgp_handle *h, mgletint level, mgletreal *u, mgletreal *v,
                        mgletreal *w, mgletint sum) {
    MGP_RANGE("mgp_par_ftoc_device");
    if (!h) return fail(MGP_ERR_STATE, "handle not initialised");
    if (!u || !v || !w) return fail(MGP_ERR_ARG, "null field");
    if (level < h->minlevel || level > h->maxlevel) return fail(MGP_ERR_ARG, "level out of range");
    try {
        const auto it = h->pf_range.find(level);
        if (it != h->pf_range.end()) {
            const double *dd[3] = {h->d_ddx.as<double>(), h->d_ddy.as<double>(),
                                   h->d_ddz.as<double>()};
            launch_par_ftoc(h->d_blocks.as<Block>(), h->d_pftask.as<ParFtocTask>() + it->second.first,
                            it->second.second, u, v, w, dd, sum != 0, h->stream);
        }
        HC(hipGetLastError());
        HC(hipStreamSynchronize(h->stream));
    } catch (const HipError &e) {
        return fail(MGP_ERR_HIP, "%s", e.msg.c_str());
    }
    return MGP_OK;
}

int mgp_flow_bc_device(mgp_handle *h, mgletint level, mgletreal *u, mgletreal *v,
                       mgletreal *w, mgletreal *p) {
    MGP_RANGE("mgp_flow_bc_device");
    if (!h) return fail(MGP_ERR_STATE, "handle not initialised");
    if (!u || !v || !w || !p) return fail(MGP_ERR_ARG, "null field");
    if (level < h->minlevel || level > h->maxlevel) return fail(MGP_ERR_ARG, "level out of range");
    try {
        if (h->level_blocks.count(level)) {
            const Topo tl = h->topo(level);
            if (h->level_par.count(level))
                launch_parent_uvw(tl, u, v, w, h->d_ubnd.as<double>(), h->d_vbnd.as<double>(),
                                  h->d_wbnd.as<double>(), h->d_bbo.as<int64_t>(), h->stream);
            if (h->level_flow.count(level))
                launch_bound_flow(tl, u, v, w, p, h->d_ubnd.as<double>(), h->d_vbnd.as<double>(),
                                  h->d_wbnd.as<double>(), h->d_bbo.as<int64_t>(),
                                  h->d_pinf.as<double>(), h->d_bpsp.as<double>(),
                                  h->d_ddx.as<double>(), h->d_ddy.as<double>(),
                                  h->d_ddz.as<double>(), h->cfg.rho, h->stream);
        }
        HC(hipGetLastError());
        HC(hipStreamSynchronize(h->stream));
    } catch (const HipError &e) {
        return fail(MGP_ERR_HIP, "%s", e.msg.c_str());
    }
    return MGP_OK;
}

// divcal into `tmp` (WO order, ncell values: every one written), then
// scattered into the interior of `div`
static void divcal_via(mgp_handle *h, double *div, const double *u, const double *v,
                       const double *w, double fak, double *tmp) {
    Topo t = h->topo(-1);
    launch_divcal<double>(t, u, v, w, h->d_rddx.as<double>(), h->d_rddy.as<double>(),
                          h->d_rddz.as<double>(), h->cfg.ibtype == 1 ? h->d_bpsp.as<double>() : nullptr,
                          h->sdiv_sp(), fak, tmp, nullptr, h->stream);
    launch_wo_to_sp<double>(t, tmp, div, h->stream);
    HC(hipGetLastError());
    HC(hipStreamSynchronize(h->stream));
}

int mgp_divcal_device(mgp_handle *h, mgletreal *div, const mgletreal *u,
                      const mgletreal *v, const mgletreal *w, mgletreal fak) {
    MGP_RANGE("mgp_divcal_device");
    if (!h) return fail(MGP_ERR_STATE, "handle not initialised");
    if (!div || !u || !v || !w) return fail(MGP_ERR_ARG, "null field");
    try {
        DBuf tmp;
        tmp.alloc((size_t)h->ncell * 8);
        divcal_via(h, div, u, v, w, fak, tmp.as<double>());
    } catch (const HipError &e) {
        return fail(MGP_ERR_HIP, "%s", e.msg.c_str());
    }
    return MGP_OK;
}

int mgp_divcal_device_noalloc(mgp_handle *h, mgletreal *div, const mgletreal *u,
                              const mgletreal *v, const mgletreal *w, mgletreal fak) {
    MGP_RANGE("mgp_divcal_device_noalloc");
    if (!h) return fail(MGP_ERR_STATE, "handle not initialised");
    if (!div || !u || !v || !w) return fail(MGP_ERR_ARG, "null field");
    if ((size_t)h->ncell > kScratchCells)
        return fail(MGP_ERR_ARG, "%lld interior cells exceed the %zu-cell scratch of "
                    "mgp_divcal_device_noalloc (use mgp_divcal_device)",
                    (long long)h->ncell, kScratchCells);
    if (!h->scratch) return fail(MGP_ERR_STATE, "scratch array not set up");
    try {
        divcal_via(h, div, u, v, w, fak, h->scratch);
    } catch (const HipError &e) {
        return fail(MGP_ERR_HIP, "%s", e.msg.c_str());
    }
    return MGP_OK;
}

int mgp_set_sdiv(mgp_handle *h, const mgletreal *sdiv) {
    MGP_RANGE("mgp_set_sdiv");
    if (!h) return fail(MGP_ERR_STATE, "handle not initialised");
    if (h->cfg.ibtype != 1)
        return fail(MGP_ERR_ARG, "SDIV enters the ghostcell divergence only (ib = ghostcell)");
    try {
        if (!sdiv) {
            h->d_sdiv.alloc(0);
            return MGP_OK;
        }
        if (!h->d_sdiv.p) h->d_sdiv.alloc((size_t)h->n3d * sizeof(double));
        HC(hipMemcpyAsync(h->d_sdiv.p, sdiv, (size_t)h->n3d * sizeof(double),
                          hipMemcpyHostToDevice, h->stream));
        HC(hipStreamSynchronize(h->stream));
    } catch (const HipError &e) {
        return fail(MGP_ERR_HIP, "%s", e.msg.c_str());
    }
    return MGP_OK;
}

int mgp_bound_layout(mgp_handle *h, int64_t *offsets, int64_t *n) {
    MGP_RANGE("mgp_bound_layout");
    if (!h) return fail(MGP_ERR_STATE, "handle not initialised");
    if (!n) return fail(MGP_ERR_ARG, "n is NULL");
    *n = h->nbb;
    if (offsets) std::copy(h->bbo.begin(), h->bbo.end(), offsets);
    return MGP_OK;
}

int mgp_set_bound(mgp_handle *h, const mgletreal *ubuf, const mgletreal *vbuf,
                  const mgletreal *wbuf, const mgletreal *pinf) {
    MGP_RANGE("mgp_set_bound");
    if (!h) return fail(MGP_ERR_STATE, "handle not initialised");
    try {
        const mgletreal *src[3] = {ubuf, vbuf, wbuf};
        DBuf *dst[3] = {&h->d_ubnd, &h->d_vbnd, &h->d_wbnd};
        for (int q = 0; q < 3; ++q) {
            if (src[q] && h->nbb > 0)
                HC(hipMemcpyAsync(dst[q]->p, src[q], (size_t)h->nbb * sizeof(double),
                                  hipMemcpyHostToDevice, h->stream));
            else if (!src[q])
                HC(hipMemsetAsync(dst[q]->p, 0, dst[q]->n, h->stream));
        }
        if (pinf)
            HC(hipMemcpyAsync(h->d_pinf.p, pinf, 6 * (size_t)h->ngrid * sizeof(double),
                              hipMemcpyHostToDevice, h->stream));
        else
            HC(hipMemsetAsync(h->d_pinf.p, 0, h->d_pinf.n, h->stream));
        HC(hipStreamSynchronize(h->stream));
    } catch (const HipError &e) {
        return fail(MGP_ERR_HIP, "%s", e.msg.c_str());
    }
    return MGP_OK;
}

// Werner-Wengle constants (init_wernerwengle, src/flow/wernerwengle_mod.F90:
// 13-29): cwb, cpo4, cpo5, cpo6, cpo8 for the wall stress next to NOS walls
static void ww_constants(double rho, double gmol, double ww[5]) {
    const double cwa = 8.3, cwb = 1.0 / 7.0;
    const double cpo1 = 1.0 - cwb, cpo2 = 1.0 + cwb;
    ww[0] = cwb;
    ww[1] = cpo2 / cwa * std::pow(gmol / rho, cwb);
    ww[2] = 0.5 * cpo1 * std::pow(cwa, cpo2 / cpo1) * std::pow(gmol / rho, cpo2);
    ww[3] = 0.5 * gmol / rho * std::pow(cwa, 2.0 / cpo1);
    ww[4] = 2.0 / cpo2;
}

int mgp_tstle4_device(mgp_handle *h, mgletreal *uo, mgletreal *vo, mgletreal *wo,
                      const mgletreal *u, const mgletreal *v, const mgletreal *w,
                      const mgletreal *p, const mgletreal *g, mgletreal gmol,
                      mgletint ilesmodel, const mgletreal *gradp) {
    MGP_RANGE("mgp_tstle4_device");
    if (!h) return fail(MGP_ERR_STATE, "handle not initialised");
    if (!uo || !vo || !wo || !u || !v || !w || !p || !g) return fail(MGP_ERR_ARG, "null field");
    try {
        MomArgs a;
        a.uo = uo; a.vo = vo; a.wo = wo;
        a.u = u; a.v = v; a.w = w; a.p = p; a.g = g;
        a.dx = h->d_dx.as<double>(); a.dy = h->d_dy.as<double>(); a.dz = h->d_dz.as<double>();
        a.ddx = h->d_ddx.as<double>(); a.ddy = h->d_ddy.as<double>(); a.ddz = h->d_ddz.as<double>();
        a.rdx = h->d_rdx.as<double>(); a.rdy = h->d_rdy.as<double>(); a.rdz = h->d_rdz.as<double>();
        a.rddx = h->d_rddx.as<double>(); a.rddy = h->d_rddy.as<double>();
        a.rddz = h->d_rddz.as<double>();
        a.rpdx = h->d_rpd[0].as<double>(); a.rpdy = h->d_rpd[1].as<double>();
        a.rpdz = h->d_rpd[2].as<double>();
        a.rrho = 1.0 / h->cfg.rho;
        a.gmol = gmol;
        a.iles = ilesmodel == 0 ? 0.0 : 1.0;
        a.rho = h->cfg.rho;
        ww_constants(h->cfg.rho, gmol, a.ww);
        for (int q = 0; q < 3; ++q) a.gp[q] = gradp ? gradp[q] : 0.0;
        launch_tstle4(h->topo(-1), a, h->stream);
        HC(hipGetLastError());
        launch_tstle4_par(h->topo(-1), a, h->d_parlist.as<int>(), h->npar, h->stream);
        HC(hipGetLastError());
        HC(hipStreamSynchronize(h->stream));
    } catch (const HipError &e) {
        return fail(MGP_ERR_HIP, "%s", e.msg.c_str());
    }
    return MGP_OK;
}

int mgp_rkstep_device(mgp_handle *h, mgletreal *p, mgletreal *dp,
                      const mgletreal *rhsp, mgletreal frhs, mgletreal dtfu) {
    MGP_RANGE("mgp_rkstep_device");
    if (!h) return fail(MGP_ERR_STATE, "handle not initialised");
    if (!p || !dp || !rhsp) return fail(MGP_ERR_ARG, "null field");
    try {
        launch_rkstep(p, dp, rhsp, h->n3d, frhs, dtfu, h->stream);
        HC(hipGetLastError());
        HC(hipStreamSynchronize(h->stream));
    } catch (const HipError &e) {
        return fail(MGP_ERR_HIP, "%s", e.msg.c_str());
    }
    return MGP_OK;
}

int mgp_time_kernels(mgp_handle *h, mgletint niter, double *ms) {
    MGP_RANGE("mgp_time_kernels");
    if (!h || !ms || niter < 1) return fail(MGP_ERR_ARG, "bad argument");
    try {
        set_plan_shapes(h->plan_shape, h->plan_shape_bwd);
        const hipStream_t s = h->stream;
        hipEvent_t ev[4];
        for (auto &e : ev) HC(hipEventCreate(&e));
        double acc[3] = {0, 0, 0};
        Outer o = h->outer();
        Outer onever = o;
        onever.done = (int *)h->d_never.p;
        const int *never = (const int *)h->d_never.p;
        const bool classic = h->cfg.ityp == 2;
        Mask m = h->mask();
        Topo tall = h->topo(-1);
        auto run = [&](auto tag) {
            using T = decltype(tag);
            Fields<T> F = h->fields<T>();
            int rd = 0;
            for (int it = 0; it < niter; ++it) {
                float t0 = 0, t1 = 0;
                HC(hipEventRecord(ev[0], s));
                launch_forward<T>(tall, F, m, rd, 1, 0, classic, never, s);
                HC(hipEventRecord(ev[1], s));
                launch_backward<T>(tall, F, 0, classic, never, s);
                HC(hipEventRecord(ev[2], s));
                HC(hipEventSynchronize(ev[2]));
                HC(hipEventElapsedTime(&t0, ev[0], ev[1]));
                HC(hipEventElapsedTime(&t1, ev[1], ev[2]));
                acc[0] += t0;
                acc[1] += t1;
                rd ^= 1;
            }
            float t2 = 0;
            HC(hipEventRecord(ev[0], s));
            launch_outer<T>(tall, F, m, onever, h->fp32() ? h->d_rhsT.as<T>() : nullptr, s);
            HC(hipEventRecord(ev[3], s));
            HC(hipEventSynchronize(ev[3]));
            HC(hipEventElapsedTime(&t2, ev[0], ev[3]));
            acc[2] = t2;
        };
        if (h->fp32()) run(float{});
        else run(double{});
        HC(hipGetLastError());
        ms[0] = acc[0] / niter;
        ms[1] = acc[1] / niter;
        ms[2] = acc[2];
        for (auto &e : ev) (void)hipEventDestroy(e);
    } catch (const HipError &e) {
        return fail(MGP_ERR_HIP, "%s", e.msg.c_str());
    }
    return MGP_OK;
}

int mgp_plan_shape(mgp_handle *h, mgletint request, mgletint *chosen,
                   double *ms) {
    MGP_RANGE("mgp_plan_shape");
    if (!h || request < -1 || request > kNumShapes) return fail(MGP_ERR_ARG, "bad argument");
    try {
        if (request < 0) {
            // report only
        } else if (request == 0) {
            if (h->fp32()) tune_shape<float>(*h);
            else tune_shape<double>(*h);
        } else {
            h->plan_shape = h->plan_shape_bwd = request;
        }
    } catch (const HipError &e) {
        return fail(MGP_ERR_HIP, "%s", e.msg.c_str());
    }
    if (chosen) *chosen = h->plan_shape;
    if (ms)
        for (int v = 0; v < kNumShapes; ++v) ms[v] = h->shape_ms[v];
    return MGP_OK;
}

int mgp_plan_shape_backward(mgp_handle *h, mgletint *chosen) {
    MGP_RANGE("mgp_plan_shape_backward");
    if (!h || !chosen) return fail(MGP_ERR_ARG, "bad argument");
    *chosen = h->plan_shape_bwd;
    return MGP_OK;
}

int mgp_get_history(mgp_handle *h, mgletreal *out, mgletint cap,
                    mgletint *count) {
    MGP_RANGE("mgp_get_history");
    if (!h || !out || !count || cap < 0) return fail(MGP_ERR_ARG, "bad argument");
    if (!h->have_solution) return fail(MGP_ERR_STATE, "no mgpoisl call yet");
    try {
        int n = 0;
        HC(hipMemcpy(&n, (char *)h->d_scal.p + 24, sizeof(int), hipMemcpyDeviceToHost));
        n = std::min(n, std::min((int)cap, std::max(1, h->cfg.nouter)));
        if (n > 0) HC(hipMemcpy(out, h->d_hist.p, n * sizeof(double), hipMemcpyDeviceToHost));
        *count = n;
    } catch (const HipError &e) {
        return fail(MGP_ERR_HIP, "%s", e.msg.c_str());
    }
    return MGP_OK;
}

int mgp_factor_classes(mgp_handle *h, mgletint *nclasses, int64_t *ncells_stored) {
    MGP_RANGE("mgp_factor_classes");
    if (!h) return fail(MGP_ERR_STATE, "handle not initialised");
    if (nclasses) *nclasses = (mgletint)h->nfclass;
    if (ncells_stored) *ncells_stored = h->nfcell;
    return MGP_OK;
}

int mgp_info(mgp_handle *h, int64_t *ncells, int64_t *device_bytes) {
    MGP_RANGE("mgp_info");
    if (!h) return fail(MGP_ERR_STATE, "handle not initialised");
    if (ncells) *ncells = h->ncell;
    if (device_bytes) {
        size_t tot = 0;
        const DBuf *bs[] = {&h->d_blocks, &h->d_shapes, &h->d_cells, &h->d_pstart, &h->d_slot,
                            &h->d_items, &h->d_ents, &h->d_zero, &h->d_cellx, &h->d_fents,
                            &h->d_ap64, &h->d_apT, &h->d_hilf, &h->d_rhs64, &h->d_rhsT, &h->d_res,
                            &h->d_dp64, &h->d_bpw, &h->d_bpf, &h->d_bpsp, &h->d_sdiv, &h->d_hface, &h->d_rhi[0],
                            &h->d_rhi[1], &h->d_rlo, &h->d_rapT, &h->d_hface2, &h->d_ft[0], &h->d_ft[1], &h->d_ft[2],
                            &h->d_ft[3], &h->d_ft[4], &h->d_ft[5], &h->d_u, &h->d_v, &h->d_w, &h->d_p, &h->d_dpsp,
                            &h->d_aps64, &h->d_apsT, &h->d_blist_rep, &h->d_parlist};
        for (auto *b : bs) tot += b->n;
        for (auto &b : h->d_f) tot += b.n;
        *device_bytes = (int64_t)tot;
    }
    return MGP_OK;
}

}  // extern "C"
