// Runtime and C ABI of libocfd_ec.so: block storage in HBM, Fortran-layout transfers, the
// per-iteration launch sequence (optionally captured in a hipGraph), residual norms and forces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ocfd_ec.h"
#include "ocfd_dev.h"
#include "ocfd_launch.h"

using namespace ocfd;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) { g_err = msg; return code; }

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) return fail(OCFD_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct HostBlock {
    int no = 0, nx = 0, ny = 0, nz = 0;
    std::vector<Sub> subs;
    Blk d{};
    std::vector<void*> allocs;
    std::vector<double> nodes[3];    // node coordinates awaiting ocfd_setup_geometry
    size_t bytes = 0;                // device memory of the block's arrays
    bool geof_ok = false;            // mixed precision: the FP32 geometry copies match the FP64 arrays
    // the block's arrays are carved out of one device allocation: a sizing pass measures it
    char* slab = nullptr;
    size_t slab_off = 0;
    bool sizing = false;
    size_t n() const { return (size_t)d.NI * d.NJ * d.NK; }
};

}  // namespace

// Launch machinery of one block list: its stream, the per-stage graphs and the device tables
// of the batched launches and of the pipelined LU-SGS sweeps.  A context owns one for its own
// blocks; a batch owns one for the blocks of all its member contexts.
struct Engine {
    hipStream_t st = nullptr;
    bool use_graph = false;
    hipGraphExec_t exec[OCFD_NSTAGES] = {};   // one captured graph per stage of the iteration
    hipGraph_t graph[OCFD_NSTAGES] = {};
    std::vector<hipEvent_t> events;           // stage boundaries of ocfd_step_timed
    int lu_schedule = OCFD_LU_PIPELINED;     // LU-SGS sweep schedule
    int lu_tj = 8;                            // tile width of the pipelined sweep
    int lu_fail = 0;                          // 1: every sweep wait that has to poll gives up
    LuPlan lu{};                              // pipelined-sweep tables (built on first use)
    int lu_nv = 0;
    std::vector<void*> lu_allocs;
    SetPlan set{};                            // batched-launch tables (built on first use)
    std::vector<void*> set_allocs;
    // plain mode: every kernel launched once per block, eagerly (no graphs), the LU-SGS sweeps on
    // 8x8 tiles; parts[b] holds block b's own tables and shares this engine's stream
    bool per_block = false;
    std::vector<Engine> parts;
    ocfd_batch* owner = nullptr;              // set for a batch's engine: its sweep storage is the batch's pool
    // the last update took DU from the pipelined sweeps' storage, which still holds it: the block
    // arrays' DU is older (flush_du copies it out)
    bool du_pending = false;
};

// one block of an engine's list, the parameters of its case and the list position of its
// case's block 1 (block numbers inside a case's boundary table are relative to it)
struct BlockRef {
    HostBlock* hb;
    const Par* p;
    int base;
};

struct ocfd_ctx {
    ocfd_params hp{};
    Par p{};
    std::vector<HostBlock> blocks;   // index block_no - 1
    char* norm_buf = nullptr;        // residual norms: block table, parts and result (ocfd_res_norms)
    size_t norm_bytes = 0;
    double* force_buf = nullptr;     // forces: per-tile parts and the 9 sums (ocfd_forces)
    size_t force_n = 0;
    int kstep = 0;
    double tt = 0.0;
    int device = 0;
    size_t bytes = 0;                // device memory of the block arrays
    Engine e;
    ocfd_batch* batch = nullptr;     // the batch the context steps through, if any
    // the context's region of its batch's sweep storage (elements), kept while it stays in the
    // batch; lu_tj is the tile width the region was laid out for (0: no region)
    long long lu_off = 0, lu_len = 0;
    int lu_tj = 0;
    std::vector<BlockRef> refs() {
        std::vector<BlockRef> r;
        for (HostBlock& hb : blocks) r.push_back(BlockRef{&hb, &p, 0});
        return r;
    }
};

struct ocfd_batch {
    Engine e;
    std::vector<ocfd_ctx*> members;
    // sweep-ordered storage of the pipelined LU-SGS sweeps: one region per member, laid out and
    // filled (sentinels, packed geometry) when the member first steps in the batch and released when
    // it leaves, so a change of membership only rebuilds the tile tables
    double* pool = nullptr;
    long long pool_cap = 0;                               // elements
    std::vector<std::pair<long long, long long>> pool_free;   // free (offset, length), by offset
    std::vector<BlockRef> refs() {
        std::vector<BlockRef> r;
        for (ocfd_ctx* c : members) {
            int base = (int)r.size();
            for (HostBlock& hb : c->blocks) r.push_back(BlockRef{&hb, &c->p, base});
        }
        return r;
    }
};


namespace {

// n bytes of the block's slab, 256-byte aligned (the sizing pass only counts)
void* carve(HostBlock& hb, size_t n) {
    const size_t at = hb.slab_off;
    hb.slab_off += (n + 255) & ~(size_t)255;
    return hb.sizing ? nullptr : hb.slab + at;
}

int dalloc(HostBlock& hb, double** p, size_t n) {
    *p = (double*)carve(hb, n * sizeof(double));
    return OCFD_OK;
}

int balloc(HostBlock& hb, uint8_t** p, size_t n) {
    *p = (uint8_t*)carve(hb, n);
    return OCFD_OK;
}

// the device arrays of one block (the same padded box for every array), carved out of the slab
int carve_block(ocfd_ctx* c, HostBlock& hb) {
    Blk& b = hb.d;
    b.nx = hb.nx; b.ny = hb.ny; b.nz = hb.nz;
    b.NI = hb.nx + 7; b.NJ = hb.ny + 7; b.NK = hb.nz + 7;
    size_t n = hb.n();
    int NV = c->p.NV, r;
    bool visc = c->p.visc == 1, turb = c->p.turb != 0;
    for (int m = 0; m < NV; m++) {
        if ((r = dalloc(hb, &b.U[m], n)) || (r = dalloc(hb, &b.Res[m], n)) || (r = dalloc(hb, &b.DU[m], n))) return r;
    }
    double** one[] = {&b.dt, &b.mu, &b.muf, &b.mut, &b.vol, &b.dtm, &b.x, &b.y, &b.z, &b.xc, &b.yc, &b.zc, &b.aux};
    for (double** q : one) if ((r = dalloc(hb, q, n))) return r;
    for (int q = 0; q < 7; q++) if ((r = dalloc(hb, &b.prim[q], n))) return r;
    for (int q = 0; q < 3; q++) if ((r = dalloc(hb, &b.Lc[q], n)) || (r = dalloc(hb, &b.Lv[q], n)) || (r = dalloc(hb, &b.S[q], n))) return r;
    for (int q = 0; q < (NV == 6 ? 3 : 1); q++) if ((r = dalloc(hb, &b.al[q], n))) return r;
    for (int q = 0; q < 9; q++) if ((r = dalloc(hb, &b.nv[q], n))) return r;
    for (int s = 1; s < 4; s++) for (int q = 0; q < 9; q++) if ((r = dalloc(hb, &b.met[s][q], n))) return r;
    if (turb) {
        for (int q = 0; q < 9; q++) if ((r = dalloc(hb, &b.met[0][q], n))) return r;
        if ((r = dalloc(hb, &b.dw, n))) return r;
    }
    if (c->p.turb == 3 && ((r = dalloc(hb, &b.kt, n)) || (r = dalloc(hb, &b.wt, n)))) return r;
    for (int dd = 0; dd < 3; dd++) for (int m = 0; m < 5; m++) if ((r = dalloc(hb, &b.F[dd][m], n))) return r;
    if (visc) for (int dd = 0; dd < 3; dd++) for (int m = 0; m < 4; m++) if ((r = dalloc(hb, &b.Fv[dd][m], n))) return r;
    if (visc && (c->p.turb == 2 || c->p.turb == 3))
        for (int dd = 0; dd < 3; dd++) for (int m = 0; m < 2; m++) if ((r = dalloc(hb, &b.Fsa[dd][m], n))) return r;
    if (c->p.mixed) {
        // FP32 primitives in half-size slots; FP32 face terms in the first half of the FP64 arrays
        for (int q = 0; q < 6; q++) {
            double* a = nullptr;
            if ((r = dalloc(hb, &a, (n + 1) / 2))) return r;
            b.primf[q] = reinterpret_cast<float*>(a);
        }
        // FP32 geometry: normals, areas, the face metrics (and the cell metrics with a turbulence
        // model), the inviscid flux's node differences
        auto falloc = [&](float** q) {
            double* a = nullptr;
            int rr = dalloc(hb, &a, (n + 1) / 2);
            *q = reinterpret_cast<float*>(a);
            return rr;
        };
        for (int q = 0; q < 9; q++) if ((r = falloc(&b.nvf[q]))) return r;
        for (int q = 0; q < 3; q++) if ((r = falloc(&b.Sf[q]))) return r;
        for (int m = turb ? 0 : 1; m < 4; m++) for (int q = 0; q < 9; q++) if ((r = falloc(&b.metf[m][q]))) return r;
        for (int dd = 0; dd < 3; dd++) for (int q = 0; q < 3; q++) if ((r = falloc(&b.t1f[dd][q]))) return r;
        for (int dd = 0; dd < 3; dd++) {
            for (int m = 0; m < 5; m++) b.Ff[dd][m] = reinterpret_cast<float*>(b.F[dd][m]);
            for (int m = 0; m < 4; m++) b.Fvf[dd][m] = reinterpret_cast<float*>(b.Fv[dd][m]);
            for (int m = 0; m < 2; m++) b.Fsaf[dd][m] = reinterpret_cast<float*>(b.Fsa[dd][m]);
        }
    }
    size_t jk = (size_t)b.NJ * b.NK, ik = (size_t)b.NI * b.NK, ij = (size_t)b.NI * b.NJ;
    if ((r = balloc(hb, &b.bcf[0], jk)) || (r = balloc(hb, &b.bcf[1], ik)) || (r = balloc(hb, &b.bcf[2], ij))) return r;
    size_t fsz[6] = {jk, ik, ij, jk, ik, ij};
    for (int f = 0; f < 6; f++) for (int q = 0; q < 3; q++) if ((r = dalloc(hb, &b.surf[f][q], fsz[f]))) return r;
    b.flags = (int*)carve(hb, 16 * sizeof(int));
    if (c->p.NV == 6) {
        size_t ncell = (size_t)(hb.nx - 1) * (hb.ny - 1) * (hb.nz - 1);
        double* m = nullptr;
        if ((r = dalloc(hb, &m, (ncell + 63) / 64))) return r;
        b.samask = reinterpret_cast<unsigned long long*>(m);
    }
    return OCFD_OK;
}

// allocate all device arrays of one block: one zeroed allocation, so a block costs its arrays'
// bytes and not one allocation granule per array (small blocks have about 130 arrays)
int alloc_block(ocfd_ctx* c, HostBlock& hb) {
    Blk& b = hb.d;
    int r;
    hb.sizing = true; hb.slab_off = 0;
    if ((r = carve_block(c, hb))) return r;
    const size_t total = hb.slab_off;
    void* q = nullptr;
    HIPCHK(hipMalloc(&q, total));
    hb.allocs.push_back(q);
    HIPCHK(hipMemset(q, 0, total));
    hb.slab = (char*)q; hb.sizing = false; hb.slab_off = 0;
    if ((r = carve_block(c, hb))) return r;
    hb.bytes = total;
    size_t n = hb.n();
    // initial values of allocate_mem_Blocks (sub_init.f90:496,507-508)
    std::vector<double> h(n, 1.0 / c->p.Re);
    HIPCHK(hipMemcpy(b.mu, h.data(), n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b.muf, h.data(), n * sizeof(double), hipMemcpyHostToDevice));
    std::fill(h.begin(), h.end(), 1.0);
    HIPCHK(hipMemcpy(b.dtm, h.data(), n * sizeof(double), hipMemcpyHostToDevice));
    return OCFD_OK;
}

// copy a Fortran array with inclusive bounds lo..hi into the padded box of a block
int put_box(const HostBlock& hb, double* dev, const double* src, const int lo[3], const int hi[3]) {
    size_t n = hb.n();
    std::vector<double> h(n, 0.0);
    const Blk& b = hb.d;
    long n0 = hi[0] - lo[0] + 1, n1 = hi[1] - lo[1] + 1;
    for (int k = lo[2]; k <= hi[2]; k++)
        for (int j = lo[1]; j <= hi[1]; j++)
            for (int i = lo[0]; i <= hi[0]; i++)
                h[IX(b, i, j, k)] = src[(i - lo[0]) + n0 * ((j - lo[1]) + n1 * (long)(k - lo[2]))];
    HIPCHK(hipMemcpy(dev, h.data(), n * sizeof(double), hipMemcpyHostToDevice));
    return OCFD_OK;
}

int get_box(const HostBlock& hb, const double* dev, std::vector<double>& h) {
    h.resize(hb.n());
    HIPCHK(hipMemcpy(h.data(), dev, hb.n() * sizeof(double), hipMemcpyDeviceToHost));
    return OCFD_OK;
}

HostBlock* find_block(ocfd_ctx* c, int no) {
    if (!c || no < 1 || no > (int)c->blocks.size() || c->blocks[no - 1].no != no) return nullptr;
    return &c->blocks[no - 1];
}

Par to_par(const ocfd_params& h) {
    Par p{};
    p.Ma = h.Ma; p.Re = h.Re; p.gamma = h.gamma; p.Cv = h.Cv; p.Cp = h.Cp; p.PrL = h.PrL; p.PrT = h.PrT;
    p.T_inf = h.T_inf; p.Twall = h.Twall; p.AoA = h.AoA; p.A_alfa = h.A_alfa; p.A_beta = h.A_beta;
    p.P_outlet = h.P_outlet; p.CFL = h.CFL; p.dtmax = h.dtmax; p.dtmin = h.dtmin; p.dt_global = h.dt_global;
    p.w_LU = h.w_LU; p.MUT_MAX = h.MUT_MAX; p.Kt_inf = h.Kt_inf; p.Wt_inf = h.Wt_inf; p.Ldmin = h.Ldmin;
    p.Ldmax = h.Ldmax; p.Lpmin = h.Lpmin; p.Lpmax = h.Lpmax; p.Lumax = h.Lumax; p.LSAmax = h.LSAmax;
    p.NV = h.NVAR; p.visc = h.If_viscous; p.turb = h.If_viscous == 1 ? h.Iflag_turbulence_model : 0;
    p.scheme = h.Iflag_Scheme; p.flux = h.Iflag_Flux; p.recon = h.IFlag_Reconstruction; p.bscheme = h.Bound_Scheme;
    p.local_dt = h.Iflag_local_dt; p.dtime_mesh = h.If_dtime_mesh; p.positivity = h.IF_Scheme_Positivity;
    p.limit_flow = h.IFLAG_LIMIT_FLOW;
    p.mixed = h.precision == 1;
    return p;
}

// boundary indicators BcI/BcJ/BcK (set_BcK, sub_init.f90:545-577) as per-row bit flags
int upload_bcflags(HostBlock& hb) {
    const Blk& b = hb.d;
    std::vector<uint8_t> fi((size_t)b.NJ * b.NK, 0), fj((size_t)b.NI * b.NK, 0), fk((size_t)b.NI * b.NJ, 0);
    for (const Sub& B : hb.subs) {
        if (B.bc < 0) continue;
        uint8_t bit = B.face <= 3 ? 1 : 2;
        if (B.face == 1 || B.face == 4) { for (int k = B.kb; k <= B.ke - 1; k++) for (int j = B.jb; j <= B.je - 1; j++) fi[JK(b, j, k)] |= bit; }
        else if (B.face == 2 || B.face == 5) { for (int k = B.kb; k <= B.ke - 1; k++) for (int i = B.ib; i <= B.ie - 1; i++) fj[IK(b, i, k)] |= bit; }
        else { for (int j = B.jb; j <= B.je - 1; j++) for (int i = B.ib; i <= B.ie - 1; i++) fk[IJ(b, i, j)] |= bit; }
    }
    HIPCHK(hipMemcpy(b.bcf[0], fi.data(), fi.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b.bcf[1], fj.data(), fj.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b.bcf[2], fk.data(), fk.size(), hipMemcpyHostToDevice));
    return OCFD_OK;
}

int check_subs(ocfd_ctx* c, const HostBlock& hb) {
    for (const Sub& B : hb.subs) {
        if (B.face < 1 || B.face > 6) return fail(OCFD_EARG, "subface with face id outside 1..6");
        if (B.ib < 1 || B.ie > hb.nx || B.jb < 1 || B.je > hb.ny || B.kb < 1 || B.ke > hb.nz || B.ib > B.ie || B.jb > B.je || B.kb > B.ke)
            return fail(OCFD_EARG, "subface range outside block " + std::to_string(hb.no));
        if (B.bc < 0) {
            if (B.nb1 < 1 || B.nb1 > (int)c->blocks.size()) return fail(OCFD_EARG, "inner subface links to unknown block");
            if (B.face1 < 1 || B.face1 > 6) return fail(OCFD_EARG, "inner subface with linked face id outside 1..6");
            int La[3] = {std::abs(B.L1), std::abs(B.L2), std::abs(B.L3)};
            for (int q = 0; q < 3; q++) if (La[q] < 1 || La[q] > 3) return fail(OCFD_EARG, "inner subface with L outside +-1..3");
        } else if ((B.bc == 5 || B.bc == 6) && c->hp.IF_InnerFlow == 1) {
            // inner flow: boundary_BC_Inflow_Turbo / boundary_BC_Outflow_Turbo (sub_boundary.f90:28-40)
            return fail(OCFD_EUNSUPPORTED, "bc " + std::to_string(B.bc) + " in the inner-flow mode (IF_InnerFlow = 1) is outside this path; block " +
                                               std::to_string(hb.no));
        } else if (!(B.bc == 2 || B.bc == 3 || B.bc == 4 || B.bc == 5 || B.bc == 6 || B.bc == 401)) {
            return fail(OCFD_EUNSUPPORTED, "The boundary condition is not supported!!! Block_no is " + std::to_string(hb.no) +
                                               " bc=" + std::to_string(B.bc));
        }
    }
    return OCFD_OK;
}

void read_subs(HostBlock& hb, int nsub, const int32_t* a) {
    hb.subs.resize(nsub);
    for (int q = 0; q < nsub; q++) {
        const int32_t* v = a + 21 * q;
        Sub& B = hb.subs[q];
        B.ib = v[0]; B.ie = v[1]; B.jb = v[2]; B.je = v[3]; B.kb = v[4]; B.ke = v[5]; B.bc = v[6]; B.face = v[7]; B.f_no = v[8];
        B.ib1 = v[9]; B.ie1 = v[10]; B.jb1 = v[11]; B.je1 = v[12]; B.kb1 = v[13]; B.ke1 = v[14]; B.nb1 = v[15]; B.face1 = v[16];
        B.f_no1 = v[17]; B.L1 = v[18]; B.L2 = v[19]; B.L3 = v[20];
    }
}

// ---- iteration -------------------------------------------------------------------------------
// The iteration is enqueued stage by stage; every launch covers all blocks of the engine's list.
// Blocks are independent inside the residual and update (the halo was filled at the end of the
// previous iteration), so the stage-major order gives the results of the reference's
// block-major loop.
void enqueue_stage(Engine& e, const std::vector<BlockRef>& refs, const Par& p0, int stage) {
    const SetPlan& P = e.set;
    if (e.per_block) {
        // plain mode: block after block; the halo copies of the whole list in one launch
        for (size_t b = 0; b < e.parts.size(); b++) {
            if (stage == OCFD_STAGE_BC_HALO) set_bc(e.parts[b].set, e.st);
            else enqueue_stage(e.parts[b], {refs[b]}, p0, stage);
        }
        if (stage == OCFD_STAGE_BC_HALO) {
            set_halo(P, e.st);
            for (Engine& q : e.parts) set_force_vt_kw(q.set, e.st);
        }
        return;
    }
    switch (stage) {
    case OCFD_STAGE_PRE:      // Set_Un is implicit (U is not modified before the update)
        set_pre(P, e.st);
        break;
    case OCFD_STAGE_INVISCID:
        set_inviscid(P, e.st);
        break;
    case OCFD_STAGE_VISCOUS_DT:
        set_viscous_dt(P, e.st);
        // the pipelined sweeps read their inputs in sweep order
        if (e.lu_schedule == OCFD_LU_PIPELINED && e.lu.fwd) launch_lusgs_pack(e.lu, p0, e.st);
        break;
    case OCFD_STAGE_LU_FORWARD:
    case OCFD_STAGE_LU_BACKWARD:
        if (e.lu_schedule != OCFD_LU_HYPERPLANE)
            launch_lusgs_pipelined(e.lu, p0, stage == OCFD_STAGE_LU_FORWARD, e.lu_schedule == OCFD_LU_PIPELINED_FENCED, e.st);
        else
            for (const BlockRef& r : refs) launch_lusgs_sweep(r.hb->d, *r.p, stage == OCFD_STAGE_LU_FORWARD, e.st);
        break;
    case OCFD_STAGE_UPDATE:
        if (e.lu_schedule == OCFD_LU_PIPELINED && e.lu.fwd) set_update_limit_sweep_order(P, e.lu, e.st);
        else set_update_limit(P, e.st);
        break;
    case OCFD_STAGE_BC_HALO:
        set_bc(P, e.st);
        set_halo(P, e.st);
        set_force_vt_kw(P, e.st);
        break;
    }
}

void drop_graphs(Engine& e) {
    for (int s = 0; s < OCFD_NSTAGES; s++) {
        if (e.exec[s]) { (void)hipGraphExecDestroy(e.exec[s]); e.exec[s] = nullptr; }
        if (e.graph[s]) { (void)hipGraphDestroy(e.graph[s]); e.graph[s] = nullptr; }
    }
}

// One stage of the iteration: eager launches, or the stage's graph (captured on first use).
int run_stage(Engine& e, const std::vector<BlockRef>& refs, const Par& p0, int stage) {
    if (stage == OCFD_STAGE_UPDATE && e.lu_schedule == OCFD_LU_PIPELINED) e.du_pending = true;
    if (!e.use_graph || e.per_block) { enqueue_stage(e, refs, p0, stage); return OCFD_OK; }
    if (!e.exec[stage]) {
        HIPCHK(hipStreamBeginCapture(e.st, hipStreamCaptureModeThreadLocal));
        enqueue_stage(e, refs, p0, stage);
        HIPCHK(hipStreamEndCapture(e.st, &e.graph[stage]));
        HIPCHK(hipGraphInstantiate(&e.exec[stage], e.graph[stage], nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(e.exec[stage], e.st));
    return OCFD_OK;
}

// The latest DU into the block arrays, from the sweep-ordered storage where the update read it
// (in plain mode from every block's own storage).  Called before the storage or its tables go
// away and when the host asks for DU.
void flush_du(Engine& e) {
    if (!e.du_pending) return;
    e.du_pending = false;
    if (e.lu_schedule != OCFD_LU_PIPELINED) return;
    if (e.per_block) {
        for (Engine& q : e.parts) if (q.lu.fwd) launch_lusgs_du_to_box(q.lu, q.lu_nv, e.st);
    } else if (e.lu.fwd) {
        launch_lusgs_du_to_box(e.lu, e.lu_nv, e.st);
    }
    (void)hipStreamSynchronize(e.st);
}

void drop_lu_plan(Engine& e) {
    flush_du(e);
    for (void* q : e.lu_allocs) (void)hipFree(q);
    e.lu_allocs.clear();
    e.lu = LuPlan{};
    e.set.halt = nullptr;
}

void drop_set_plan(Engine& e) {
    for (void* q : e.set_allocs) (void)hipFree(q);
    e.set_allocs.clear();
    e.set = SetPlan{};
}

// every table of the engine, e.g. when its block list changes
void drop_plans(Engine& e) {
    drop_graphs(e);
    drop_lu_plan(e);
    drop_set_plan(e);
    for (Engine& q : e.parts) drop_plans(q);      // the parts share the stream: nothing else to free
    e.parts.clear();
}

template <typename T>
int upload_table(std::vector<void*>& allocs, const std::vector<T>& v, T** out) {
    void* q = nullptr;
    size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
    HIPCHK(hipMalloc(&q, bytes));
    allocs.push_back(q);
    if (!v.empty()) HIPCHK(hipMemcpy(q, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    *out = (T*)q;
    return OCFD_OK;
}

// Device tables of the pipelined sweeps: the block table, the tile lists of both sweep
// directions, one progress counter per tile and direction (-1 = nothing done), the error flag
// and the sweep-ordered storage with the geometry packed into it.
int build_lu_plan_batch(ocfd_batch* b, const Par& p0);

int build_lu_plan(Engine& e, const std::vector<BlockRef>& refs, const Par& p0) {
    const int NV = p0.NV;
    if (e.lu.fwd) return OCFD_OK;
    if (e.owner) return build_lu_plan_batch(e.owner, p0);
    std::vector<Blk> blks;
    for (const BlockRef& r : refs) blks.push_back(r.hb->d);
    long long P = 0;
    const int tj = e.lu_schedule == OCFD_LU_PIPELINED_FENCED ? 8 : e.lu_tj;
    std::vector<LuTask> f = lusgs_tasks(blks, true, lusgs_fields(NV), tj, &P), b = lusgs_tasks(blks, false, lusgs_fields(NV), tj, nullptr);
    std::vector<int> minus(std::max(f.size(), b.size()), -1), zero{0, 0, e.lu_fail, 0};
    std::vector<int> mf(minus.begin(), minus.begin() + f.size()), mb(minus.begin(), minus.begin() + b.size());
    Blk* pb; LuTask *pf, *pw; int *qf, *qb, *pe;
    int r;
    if ((r = upload_table(e.lu_allocs, blks, &pb)) || (r = upload_table(e.lu_allocs, f, &pf)) ||
        (r = upload_table(e.lu_allocs, b, &pw)) || (r = upload_table(e.lu_allocs, mf, &qf)) ||
        (r = upload_table(e.lu_allocs, mb, &qb)) || (r = upload_table(e.lu_allocs, zero, &pe)))
        return r;
    e.lu.blks = pb; e.lu.fwd = pf; e.lu.bwd = pw;
    e.lu.n_fwd = (int)f.size(); e.lu.n_bwd = (int)b.size();
    e.lu.prog_fwd = qf; e.lu.prog_bwd = qb; e.lu.err = pe;
    void* pk = nullptr;
    size_t bytes = (size_t)P * sizeof(double);
    HIPCHK(hipMalloc(&pk, bytes));
    e.lu_allocs.push_back(pk);
    // on the engine's stream: a plain hipMemset runs on the null stream, which does not order
    // against a non-blocking stream, and could land after the geometry packing below
    HIPCHK(hipMemsetAsync(pk, 0, bytes, e.st));
    e.lu.pk = (double*)pk;
    e.lu.P = P;
    const std::vector<int2> wins = lusgs_window_list(blks, f);
    int2* pwin;
    if ((r = upload_table(e.lu_allocs, wins, &pwin))) return r;
    e.lu.win = pwin;
    e.lu.nwins = (int)wins.size();
    e.lu.tj = tj;
    launch_lusgs_pack_geometry(e.lu, p0, e.st);
    launch_lusgs_reset(e.lu, NV, e.st);
    e.lu_nv = NV;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e.st));
    return OCFD_OK;
}

// ---- a batch's sweep storage pool ----
// Regions are multiples of one step of a tile (NF x 64 elements), so a member's storage keeps the
// alignment of a context's own (the fields of an element follow from its offset).
void pool_release(ocfd_batch* b, ocfd_ctx* c) {
    if (c->lu_tj == 0) return;
    auto& fr = b->pool_free;
    fr.emplace_back(c->lu_off, c->lu_len);
    std::sort(fr.begin(), fr.end());
    std::vector<std::pair<long long, long long>> merged;
    for (auto& r : fr) {
        if (!merged.empty() && merged.back().first + merged.back().second == r.first) merged.back().second += r.second;
        else merged.push_back(r);
    }
    fr.swap(merged);
    c->lu_tj = 0; c->lu_off = c->lu_len = 0;
}

void pool_destroy(ocfd_batch* b) {
    for (ocfd_ctx* c : b->members) { c->lu_tj = 0; c->lu_off = c->lu_len = 0; }
    if (b->pool) (void)hipFree(b->pool);
    b->pool = nullptr; b->pool_cap = 0; b->pool_free.clear();
}

// a free range of n elements (first fit); the pool grows (contents copied) when none is left
int pool_take(ocfd_batch* b, long long n, long long step, long long* off) {
    auto& fr = b->pool_free;
    for (size_t q = 0; q < fr.size(); q++) {
        if (fr[q].second < n) continue;
        *off = fr[q].first;
        fr[q].first += n; fr[q].second -= n;
        if (fr[q].second == 0) fr.erase(fr.begin() + q);
        return OCFD_OK;
    }
    const long long cap = (std::max(b->pool_cap + n, b->pool_cap + b->pool_cap / 2) + step - 1) / step * step;
    void* q = nullptr;
    HIPCHK(hipMalloc(&q, (size_t)cap * sizeof(double)));
    if (b->pool) {
        HIPCHK(hipMemcpyAsync(q, b->pool, (size_t)b->pool_cap * sizeof(double), hipMemcpyDeviceToDevice, b->e.st));
        HIPCHK(hipStreamSynchronize(b->e.st));
        HIPCHK(hipFree(b->pool));
    }
    // the old tail (free or not) plus the new space, merged with a free range ending at the old cap
    long long start = b->pool_cap;
    if (!fr.empty() && fr.back().first + fr.back().second == b->pool_cap) { start = fr.back().first; fr.pop_back(); }
    b->pool = (double*)q;
    b->pool_cap = cap;
    *off = start;
    if (start + n < cap) fr.emplace_back(start + n, cap - start - n);
    return OCFD_OK;
}

// A member's region: its tiles laid out as a context's own sweep storage would be, zeroed, the
// sentinel in its DU fields and padding, its geometry packed (through a temporary table set)
int pool_fill_member(ocfd_batch* b, ocfd_ctx* c, int tj, const Par& p0) {
    const int NV = p0.NV;
    std::vector<Blk> blks;
    for (HostBlock& hb : c->blocks) blks.push_back(hb.d);
    long long P = 0;
    std::vector<LuTask> f = lusgs_tasks(blks, true, lusgs_fields(NV), tj, &P);
    const long long step = (long long)lusgs_fields(NV) * 64;
    P = (P + step - 1) / step * step;
    long long off = 0;
    int r = pool_take(b, P, step, &off);
    if (r) return r;
    c->lu_off = off; c->lu_len = P; c->lu_tj = tj;
    std::vector<void*> tmp;
    const std::vector<int2> wins = lusgs_window_list(blks, f);
    Blk* pb; LuTask* pf; int2* pwin;
    if ((r = upload_table(tmp, blks, &pb)) || (r = upload_table(tmp, f, &pf)) || (r = upload_table(tmp, wins, &pwin))) {
        for (void* q : tmp) (void)hipFree(q);
        return r;
    }
    LuPlan L{};
    L.blks = pb; L.fwd = pf; L.n_fwd = (int)f.size();
    L.pk = b->pool + off; L.P = P; L.tj = tj;
    L.win = pwin; L.nwins = (int)wins.size();
    HIPCHK(hipMemsetAsync(L.pk, 0, (size_t)P * sizeof(double), b->e.st));
    launch_lusgs_pack_geometry(L, p0, b->e.st);
    launch_lusgs_reset(L, NV, b->e.st);
    hipError_t e = hipStreamSynchronize(b->e.st);
    for (void* q : tmp) (void)hipFree(q);
    if (e != hipSuccess) return fail(OCFD_EHIP, std::string("filling a member's sweep storage: ") + hipGetErrorString(e));
    return OCFD_OK;
}

// The batch's sweep tables over the members' regions: each member's tiles as it would plan them
// alone, with storage offsets moved to its region and block and tile numbers to its place in the
// batch's lists.  A member without a region (new, or laid out for another tile width) gets one.
int build_lu_plan_batch(ocfd_batch* b, const Par& p0) {
    Engine& e = b->e;
    const int NV = p0.NV;
    const int tj = e.lu_schedule == OCFD_LU_PIPELINED_FENCED ? 8 : e.lu_tj;
    int r;
    for (ocfd_ctx* c : b->members)
        if (c->lu_tj != 0 && c->lu_tj != tj) pool_release(b, c);
    for (ocfd_ctx* c : b->members)
        if (c->lu_tj == 0 && (r = pool_fill_member(b, c, tj, p0))) return r;
    std::vector<Blk> blks;
    std::vector<LuTask> f, w;
    for (ocfd_ctx* c : b->members) {
        std::vector<Blk> mb;
        for (HostBlock& hb : c->blocks) mb.push_back(hb.d);
        const int bb = (int)blks.size(), tf = (int)f.size(), tw = (int)w.size();
        std::vector<LuTask> mf = lusgs_tasks(mb, true, lusgs_fields(NV), tj, nullptr);
        std::vector<LuTask> mw = lusgs_tasks(mb, false, lusgs_fields(NV), tj, nullptr);
        auto shift = [&](std::vector<LuTask>& v, int tb, std::vector<LuTask>& out) {
            for (LuTask t : v) {
                t.blk += bb;
                t.pofs += c->lu_off; t.pofs_j += c->lu_off; t.pofs_k += c->lu_off;
                if (t.dep_j >= 0) t.dep_j += tb;
                if (t.dep_k >= 0) t.dep_k += tb;
                out.push_back(t);
            }
        };
        shift(mf, tf, f);
        shift(mw, tw, w);
        for (const Blk& q : mb) blks.push_back(q);
    }
    std::vector<int> mf(f.size(), -1), mb(w.size(), -1), zero{0, 0, e.lu_fail, 0};
    Blk* pb; LuTask *pf, *pw; int *qf, *qb, *pe;
    if ((r = upload_table(e.lu_allocs, blks, &pb)) || (r = upload_table(e.lu_allocs, f, &pf)) ||
        (r = upload_table(e.lu_allocs, w, &pw)) || (r = upload_table(e.lu_allocs, mf, &qf)) ||
        (r = upload_table(e.lu_allocs, mb, &qb)) || (r = upload_table(e.lu_allocs, zero, &pe)))
        return r;
    e.lu.blks = pb; e.lu.fwd = pf; e.lu.bwd = pw;
    e.lu.n_fwd = (int)f.size(); e.lu.n_bwd = (int)w.size();
    e.lu.prog_fwd = qf; e.lu.prog_bwd = qb; e.lu.err = pe;
    const std::vector<int2> wins = lusgs_window_list(blks, f);
    int2* pwin;
    if ((r = upload_table(e.lu_allocs, wins, &pwin))) return r;
    e.lu.pk = b->pool;
    e.lu.P = b->pool_cap;
    e.lu.win = pwin;
    e.lu.nwins = (int)wins.size();
    e.lu.tj = tj;
    e.lu_nv = NV;
    return OCFD_OK;
}

// After a sweep gave up waiting: clear the flag and re-arm both counter sets.
int check_lu_error(Engine& e) {
    if (!e.lu.err) return OCFD_OK;
    int v = 0;
    HIPCHK(hipMemcpy(&v, e.lu.err, sizeof(int), hipMemcpyDeviceToHost));
    if (v == 0) return OCFD_OK;
    std::vector<int> minus(std::max(e.lu.n_fwd, e.lu.n_bwd), -1);
    HIPCHK(hipMemcpy(e.lu.prog_fwd, minus.data(), e.lu.n_fwd * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e.lu.prog_bwd, minus.data(), e.lu.n_bwd * sizeof(int), hipMemcpyHostToDevice));
    int z = 0;
    HIPCHK(hipMemcpy(e.lu.err, &z, sizeof(int), hipMemcpyHostToDevice));
    // the storage's DU is not an iteration's any more
    e.du_pending = false;
    if (e.owner) {
        // a batch: every member's region as laid out on its own
        for (ocfd_ctx* c : e.owner->members) {
            if (c->lu_tj == 0) continue;
            LuPlan L{};
            L.pk = e.owner->pool + c->lu_off; L.P = c->lu_len;
            launch_lusgs_reset(L, e.lu_nv, e.st);
        }
    } else {
        launch_lusgs_reset(e.lu, e.lu_nv, e.st);
    }
    HIPCHK(hipStreamSynchronize(e.st));
    return fail(OCFD_EHIP, "LU-SGS sweep: a tile waited beyond its poll limit; the iteration is invalid");
}

// extent of a subface launch (the kernels check the exact inclusive/exclusive node ranges)
void sub_extent(const Sub& B, int& n0, int& n1) {
    switch (B.face) {
    case 1: case 4: n0 = B.je - B.jb + 1; n1 = B.ke - B.kb + 1; break;
    case 2: case 5: n0 = B.ie - B.ib + 1; n1 = B.ke - B.kb + 1; break;
    default: n0 = B.ie - B.ib + 1; n1 = B.je - B.jb + 1; break;
    }
}

int build_sub_list(std::vector<void*>& allocs, std::vector<SubTask>& tasks, SubList& L) {
    std::vector<int> tile_task;
    for (size_t q = 0; q < tasks.size(); q++) {
        SubTask& t = tasks[q];
        t.tile0 = (int)tile_task.size();
        int nt = ((t.n0 + 63) / 64) * std::max(t.n1, 0);
        for (int a = 0; a < nt; a++) tile_task.push_back((int)q);
    }
    L.ntiles = (int)tile_task.size();
    int r;
    if ((r = upload_table(allocs, tasks, &L.t)) || (r = upload_table(allocs, tile_task, &L.tile_task))) return r;
    return OCFD_OK;
}

// Device tables of the batched launches: blocks, their parameters, the tile maps of every
// iteration space and the task lists of the subface launches.  Built once per block list.
int build_set_plan(Engine& e, const std::vector<BlockRef>& refs, const Par& p0) {
    if (e.set.blk) return OCFD_OK;
    SetPlan& P = e.set;
    std::vector<Blk> blks;
    std::vector<Par> pars;
    for (const BlockRef& r : refs) { blks.push_back(r.hb->d); pars.push_back(*r.p); }
    P.nb = (int)blks.size();
    P.p0 = p0;
    int r;
    if ((r = upload_table(e.set_allocs, blks, &P.blk)) || (r = upload_table(e.set_allocs, pars, &P.par))) return r;
    for (int sp = 0; sp < SP_COUNT; sp++) {
        std::vector<int> tb, t0;
        for (int bi = 0; bi < P.nb; bi++) {
            int n0, n1, n2, o0, o1, o2;
            space_dims(sp, blks[bi], n0, n1, n2, o0, o1, o2);
            t0.push_back((int)tb.size());
            long nt = ((long)n0 * n1 * n2 + 63) / 64;
            for (long a = 0; a < nt; a++) tb.push_back(bi);
        }
        P.ntiles[sp] = (int)tb.size();
        if ((r = upload_table(e.set_allocs, tb, &P.tile_blk[sp])) || (r = upload_table(e.set_allocs, t0, &P.tile0[sp]))) return r;
    }
    std::vector<SubTask> bc, amut, kbc, force;
    std::vector<HaloTask> halo;
    std::vector<int> halo_tile;
    for (int bi = 0; bi < P.nb; bi++) {
        const HostBlock& hb = *refs[bi].hb;
        const Par& p = *refs[bi].p;
        P.mx = std::max(P.mx, std::max(hb.nx, std::max(hb.ny, hb.nz)));
        for (const Sub& B : hb.subs) {
            SubTask t{};
            t.blk = bi; t.B = B;
            sub_extent(B, t.n0, t.n1);
            if (B.bc > 0) {
                if (B.bc == 2 && p.visc == 1) t.kind = SUB_WALL;
                else if (B.bc == 4) { t.kind = SUB_FAR; t.arg = 0; }
                else if (B.bc == 5) { t.kind = SUB_FAR; t.arg = 2; }
                else if (B.bc == 6) { t.kind = SUB_FAR; t.arg = 1; }
                else if (B.bc == 3 || B.bc == 2) t.kind = SUB_SYM;
                else t.kind = SUB_EXT;
                bc.push_back(t);
                if (B.bc == 2) {
                    SubTask a = t; a.kind = SUB_AMUT; amut.push_back(a);
                    // comput_force's face loop: cells of the two tangential directions
                    SubTask fo = t;
                    const int d = (B.face - 1) % 3;
                    fo.n0 = d == 0 ? B.je - B.jb : B.ie - B.ib;
                    fo.n1 = d == 2 ? B.je - B.jb : B.ke - B.kb;
                    if (fo.n0 > 0 && fo.n1 > 0) force.push_back(fo);
                }
                if (B.face == 3 || B.face == 6) {
                    SubTask k = t;
                    k.kind = SUB_KBC; k.n0 = B.ie - B.ib; k.n1 = std::max(1, B.je - B.jb);
                    kbc.push_back(k);
                }
            } else if (B.bc < 0) {
                HaloTask h{};
                h.src = bi; h.dst = refs[bi].base + B.nb1 - 1; h.h = halo_map(B); h.n = halo_points(h.h);
                h.tile0 = (int)halo_tile.size();
                for (int a = 0; a < (h.n + 255) / 256; a++) halo_tile.push_back((int)halo.size());
                halo.push_back(h);
            }
        }
    }
    std::vector<SubTask> vghost;
    for (int bi = 0; bi < P.nb; bi++) {
        const Blk& b = blks[bi];
        for (int f = 0; f < 6; f++) {
            SubTask t{};
            t.blk = bi; t.arg = f;
            t.n0 = f < 2 ? b.ny - 1 : b.nx - 1;
            t.n1 = f < 4 ? b.nz - 1 : b.ny - 1;
            vghost.push_back(t);
        }
    }
    if ((r = build_sub_list(e.set_allocs, bc, P.bc)) || (r = build_sub_list(e.set_allocs, amut, P.amut)) ||
        (r = build_sub_list(e.set_allocs, kbc, P.kbc)) || (r = build_sub_list(e.set_allocs, force, P.force)) ||
        (r = build_sub_list(e.set_allocs, vghost, P.vghost)))
        return r;
    P.halo.ntiles = (int)halo_tile.size();
    if ((r = upload_table(e.set_allocs, halo, &P.halo.t)) || (r = upload_table(e.set_allocs, halo_tile, &P.halo.tile_task))) return r;
    return OCFD_OK;
}

// the tables an iteration's launches read (built outside any graph capture)
int engine_prepare(Engine& e, const std::vector<BlockRef>& refs, const Par& p0, bool lu) {
    int r = build_set_plan(e, refs, p0);
    if (r) return r;
    if (p0.mixed) {
        for (const BlockRef& q : refs)
            if (!q.hb->geof_ok) { launch_geometry_f32(q.hb->d, e.st); q.hb->geof_ok = true; }
        HIPCHK(hipGetLastError());
    }
    if (e.per_block) {
        if (e.parts.size() != refs.size()) {
            for (Engine& q : e.parts) drop_plans(q);
            e.parts.clear();
            e.parts.resize(refs.size());
            for (Engine& q : e.parts) { q.st = e.st; q.lu_schedule = e.lu_schedule; q.lu_tj = 8; }
        }
        for (size_t b = 0; b < refs.size(); b++) {
            Engine& q = e.parts[b];
            q.lu_fail = e.lu_fail;
            // a one-block list: its inner links point outside it, and the halo runs on the whole list
            BlockRef one = refs[b];
            one.base = 0;
            if ((r = engine_prepare(q, {one}, p0, lu))) return r;
        }
        return OCFD_OK;
    }
    if (lu && e.lu_schedule != OCFD_LU_HYPERPLANE) {
        if ((r = build_lu_plan(e, refs, p0))) return r;
        // once a sweep gives up, the update kernels of that and later iterations leave U as it is
        e.set.halt = e.lu.err;
    }
    return OCFD_OK;
}

int engine_finish(Engine& e) {
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(e.st));
    int r = OCFD_OK;
    for (Engine& q : e.parts) { int rq = check_lu_error(q); if (rq && !r) r = rq; }
    if (r) { e.du_pending = false; return r; }     // the parts' storage holds no iteration's DU
    return check_lu_error(e);
}

// niter iterations; with stage_ms, HIP events at the stage boundaries (device ms per stage,
// summed over the iterations)
int engine_step(Engine& e, const std::vector<BlockRef>& refs, const Par& p0, int niter, double* stage_ms) {
    int r;
    if (stage_ms) {
        size_t need = (size_t)niter * (OCFD_NSTAGES + 1);
        while (e.events.size() < need) {
            hipEvent_t ev;
            HIPCHK(hipEventCreate(&ev));
            e.events.push_back(ev);
        }
    }
    for (int it = 0; it < niter; it++) {
        hipEvent_t* ev = stage_ms ? e.events.data() + (size_t)it * (OCFD_NSTAGES + 1) : nullptr;
        if (ev) HIPCHK(hipEventRecord(ev[0], e.st));
        for (int s = 0; s < OCFD_NSTAGES; s++) {
            if ((r = run_stage(e, refs, p0, s))) return r;
            if (ev) HIPCHK(hipEventRecord(ev[s + 1], e.st));
        }
    }
    if ((r = engine_finish(e))) return r;
    if (stage_ms) {
        for (int s = 0; s < OCFD_NSTAGES; s++) stage_ms[s] = 0.0;
        for (int it = 0; it < niter; it++) {
            hipEvent_t* ev = e.events.data() + (size_t)it * (OCFD_NSTAGES + 1);
            for (int s = 0; s < OCFD_NSTAGES; s++) {
                float ms = 0.0f;
                HIPCHK(hipEventElapsedTime(&ms, ev[s], ev[s + 1]));
                stage_ms[s] += (double)ms;
            }
        }
    }
    return OCFD_OK;
}

void drop_parts(Engine& e) {
    for (Engine& q : e.parts) drop_plans(q);
    e.parts.clear();
}

int engine_set_lu_tile(Engine& e, int tj) {
    if (!lusgs_tile_ok(tj)) return fail(OCFD_EARG, "LU-SGS tile width must be 4, 8 or 16");
    if (e.per_block && tj != 8) return fail(OCFD_EARG, "the per-block launch mode keeps 8x8 tiles");
    if (tj != e.lu_tj) { drop_graphs(e); drop_lu_plan(e); drop_parts(e); e.lu_tj = tj; }
    return OCFD_OK;
}

int engine_set_lu_schedule(Engine& e, int schedule) {
    if (schedule != OCFD_LU_PIPELINED && schedule != OCFD_LU_HYPERPLANE && schedule != OCFD_LU_PIPELINED_FENCED)
        return fail(OCFD_EARG, "unknown LU-SGS schedule");
    // the fenced sweep plans 8x8 tiles, the team sweep the engine's tile width
    if (schedule != e.lu_schedule) { drop_graphs(e); drop_lu_plan(e); drop_parts(e); e.lu_schedule = schedule; }
    return OCFD_OK;
}

// The LU-SGS tile-shape autotuner: niter ordinary iterations with each tile shape (4x16, 8x8,
// 16x4 (j,k) columns), timing the two sweeps of each with HIP events, then the fastest shape is
// kept.  Every shape gives the same result bit for bit, so the solve advances exactly as by
// 3 niter plain iterations.
int engine_autotune(Engine& e, const std::vector<BlockRef>& refs, const Par& p0, int niter, int* best, double* ms3) {
    if (e.lu_schedule != OCFD_LU_PIPELINED) return fail(OCFD_EARG, "tile shapes apply to the pipelined LU-SGS schedule");
    if (e.per_block) return fail(OCFD_EARG, "the per-block launch mode keeps 8x8 tiles");
    if (niter < 1) return fail(OCFD_EARG, "niter < 1");
    static const int shapes[3] = {4, 8, 16};
    double t[3];
    int r;
    for (int q = 0; q < 3; q++) {
        if ((r = engine_set_lu_tile(e, shapes[q])) || (r = engine_prepare(e, refs, p0, true))) return r;
        double sm[OCFD_NSTAGES];
        if ((r = engine_step(e, refs, p0, niter, sm))) return r;
        t[q] = (sm[OCFD_STAGE_LU_FORWARD] + sm[OCFD_STAGE_LU_BACKWARD]) / niter;
        if (ms3) ms3[q] = t[q];
    }
    int k = 0;
    for (int q = 1; q < 3; q++) if (t[q] < t[k]) k = q;
    if ((r = engine_set_lu_tile(e, shapes[k]))) return r;
    if (best) *best = shapes[k];
    return OCFD_OK;
}

void engine_destroy(Engine& e) {
    if (e.st) (void)hipStreamSynchronize(e.st);
    drop_plans(e);
    for (hipEvent_t ev : e.events) (void)hipEventDestroy(ev);
    e.events.clear();
    if (e.st) (void)hipStreamDestroy(e.st);
    e.st = nullptr;
}

int ready(ocfd_ctx* c) {
    if (!c) return fail(OCFD_EARG, "null context");
    if (c->blocks.empty()) return fail(OCFD_EARG, "no blocks added");
    for (size_t q = 0; q < c->blocks.size(); q++)
        if (c->blocks[q].no != (int)q + 1) return fail(OCFD_EARG, "block " + std::to_string(q + 1) + " missing");
    return build_set_plan(c->e, c->refs(), c->p);
}

// ready() for the entry points that advance or change the state: a context in a batch is stepped
// through the batch only
int ready_solo(ocfd_ctx* c) {
    if (c && c->batch) return fail(OCFD_EARG, "the context is in a batch and is stepped through it (ocfd_batch_remove first)");
    return ready(c);
}

// ready_solo() plus the LU-SGS tables
int prepare(ocfd_ctx* c) {
    int r = ready_solo(c);
    if (r) return r;
    return engine_prepare(c->e, c->refs(), c->p, true);
}

int finish(ocfd_ctx* c) { return engine_finish(c->e); }

void enqueue_stage(ocfd_ctx* c, int stage) {
    if (stage == OCFD_STAGE_UPDATE && c->e.lu_schedule == OCFD_LU_PIPELINED) c->e.du_pending = true;
    enqueue_stage(c->e, c->refs(), c->p, stage);
}

// device buffer of the residual norms: the block table, the parts, the result (3 NV doubles)
struct NormLayout { size_t tab, part, out, bytes; };
NormLayout norm_layout(int nb, int NV) {
    NormLayout L;
    L.tab = 0;
    L.part = ((size_t)nb * sizeof(NormBlk) + 255) & ~(size_t)255;
    L.out = L.part + (((size_t)nb * NV * norms_parts() * 3 * sizeof(double) + 255) & ~(size_t)255);
    L.bytes = L.out + 3 * MAXV * sizeof(double);
    return L;
}

void fill_norm_table(const ocfd_ctx* c, NormBlk* tab) {
    for (size_t m = 0; m < c->blocks.size(); m++) {
        const Blk& b = c->blocks[m].d;
        NormBlk t{};
        for (int v = 0; v < MAXV; v++) t.res[v] = v < c->p.NV ? b.Res[v] : nullptr;
        t.nx = b.nx; t.ny = b.ny; t.nz = b.nz; t.NI = b.NI; t.NJ = b.NJ;
        tab[m] = t;
    }
}

// the norms' launches on the context's stream over the buffer `dev` (table already in place),
// then Res_max and Res_rms = sqrt(sum / Num_Cell) on the host (comput_max_Res_onemesh,
// sub_time_advance.f90:330-380)
int run_norms(ocfd_ctx* c, char* dev, const NormLayout& L, double* res_max, double* res_rms) {
    const int NV = c->p.NV, nb = (int)c->blocks.size();
    launch_norms(reinterpret_cast<const NormBlk*>(dev + L.tab), nb, NV, reinterpret_cast<double*>(dev + L.part),
                 reinterpret_cast<double*>(dev + L.out), c->e.st);
    HIPCHK(hipGetLastError());
    double o[3 * MAXV];
    HIPCHK(hipMemcpyAsync(o, dev + L.out, 3 * (size_t)NV * sizeof(double), hipMemcpyDeviceToHost, c->e.st));
    HIPCHK(hipStreamSynchronize(c->e.st));
    long ncell = 0;
    for (const HostBlock& hb : c->blocks) ncell += (long)(hb.nx - 1) * (hb.ny - 1) * (hb.nz - 1);
    bool nan = false;
    for (int v = 0; v < NV; v++) {
        res_max[v] = o[v];
        res_rms[v] = std::sqrt(o[NV + v] / (double)ncell);
        nan = nan || o[2 * NV + v] != 0.0;
    }
    if (nan) return fail(OCFD_ENAN, "NaN in Residual is found !");
    return OCFD_OK;
}
void enqueue_residual(ocfd_ctx* c) {
    for (int s = OCFD_STAGE_PRE; s <= OCFD_STAGE_LU_BACKWARD; s++) enqueue_stage(c, s);
    // du_LU_SGS leaves DU in the block arrays; the pipelined sweeps leave it in their storage,
    // where the update reads it, so it is copied out for the host here
    Engine& e = c->e;
    if (e.lu_schedule == OCFD_LU_PIPELINED) {
        if (e.per_block) {
            for (Engine& q : e.parts) if (q.lu.fwd) launch_lusgs_du_to_box(q.lu, c->p.NV, e.st);
        } else if (e.lu.fwd) {
            launch_lusgs_du_to_box(e.lu, c->p.NV, e.st);
        }
    }
    e.du_pending = false;
}
void enqueue_bc(ocfd_ctx* c) { set_bc(c->e.set, c->e.st); }
void enqueue_halo(ocfd_ctx* c) { set_halo(c->e.set, c->e.st); }
void enqueue_force_vt_kw(ocfd_ctx* c) { set_force_vt_kw(c->e.set, c->e.st); }

}  // namespace

namespace ocfd {

// The residual norms of a context computed in a caller-provided device buffer of cap bytes (the
// block table is copied in from the host stack, so the buffer may hold anything before): no
// allocation of any kind on the device.
int norms_into(ocfd_ctx* c, char* dev, size_t cap, double* res_max, double* res_rms) {
    if (!c) return fail(OCFD_EARG, "null context");
    if (!res_max || !res_rms) return fail(OCFD_EARG, "null output");
    const int nb = (int)c->blocks.size();
    if (nb < 1) return fail(OCFD_EARG, "no blocks added");
    for (int q = 0; q < nb; q++)
        if (c->blocks[q].no != q + 1) return fail(OCFD_EARG, "block " + std::to_string(q + 1) + " missing");
    if (nb > NORMS_INTO_MAX_BLOCKS)
        return fail(OCFD_EARG, "at most " + std::to_string(NORMS_INTO_MAX_BLOCKS) + " blocks; this context has " + std::to_string(nb));
    const NormLayout L = norm_layout(nb, c->p.NV);
    if (L.bytes > cap) return fail(OCFD_EARG, "the residual norms of this context exceed the buffer");
    HIPCHK(hipSetDevice(c->device));
    NormBlk tab[NORMS_INTO_MAX_BLOCKS];
    fill_norm_table(c, tab);
    HIPCHK(hipMemcpyAsync(dev + L.tab, tab, (size_t)nb * sizeof(NormBlk), hipMemcpyHostToDevice, c->e.st));
    return run_norms(c, dev, L, res_max, res_rms);
}

// ---- the device scratch of ocfd_res_norms_noalloc ----
// The norms' block table, per-workgroup parts and sums for a context of at most
// NORMS_INTO_MAX_BLOCKS blocks, as a file-scope device array: every GPU the library runs on holds
// its own copy (about 0.7 MB) until the process ends.  In host code the array's name is not a
// device address; its address on a device comes from hipGetSymbolAddress, taken once per device
// with that device current (ocfd_create) and kept in g_scratch.
constexpr size_t SCRATCH_BYTES = (((size_t)NORMS_INTO_MAX_BLOCKS * sizeof(NormBlk) + 255) & ~(size_t)255) +
                                 (((size_t)NORMS_INTO_MAX_BLOCKS * MAXV * NORMS_PARTS * 3 * sizeof(double) + 255) & ~(size_t)255) +
                                 3 * MAXV * sizeof(double);
static __device__ __align__(256) char scratch[SCRATCH_BYTES];
constexpr int SCRATCH_MAX_DEVICES = 64;
static char* g_scratch[SCRATCH_MAX_DEVICES];
static std::mutex g_scratch_mu;

// the scratch address on the current device `dev` (once per device)
int scratch_setup(int dev) {
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return fail(OCFD_EHIP, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[dev]) return OCFD_OK;
    void* q = nullptr;
    const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !q) return fail(OCFD_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = static_cast<char*>(q);
    return OCFD_OK;
}

char* scratch_of(int dev) {
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch[dev];
}

int ctx_device(const ocfd_ctx* c) { return c ? c->device : -1; }
int set_error(int code, const char* msg) { return fail(code, msg); }

}  // namespace ocfd

// ================================ C ABI ========================================================
extern "C" {

const char* ocfd_last_error(void) { return g_err.c_str(); }

int ocfd_create(const ocfd_params* p, ocfd_ctx** out) {
    if (!p || !out) return fail(OCFD_EARG, "null argument");
    if (p->NVAR < 5 || p->NVAR > 7) return fail(OCFD_EARG, "NVAR must be 5, 6 or 7");
    if (p->If_viscous == 1 && !(p->Iflag_turbulence_model == 0 || p->Iflag_turbulence_model == 2 || p->Iflag_turbulence_model == 3))
        return fail(OCFD_EUNSUPPORTED, "turbulence model outside SA (2) and SST (3)");
    if (p->Iflag_Scheme == 9 || p->Bound_Scheme == 9) return fail(OCFD_EUNSUPPORTED, "WENO7 reconstruction is not on this path");
    if (p->Iflag_Flux < 1 || p->Iflag_Flux > 6) return fail(OCFD_EARG, "Iflag_Flux outside 1..6");
    if (p->precision != 0 && p->precision != 1) return fail(OCFD_EARG, "precision must be 0 (all FP64) or 1 (FP32 reconstruction, fluxes and SA terms)");
    if (p->IF_TurboMachinary != 0) return fail(OCFD_EUNSUPPORTED, "the turbomachinery mode (IF_TurboMachinary = 1) is outside this path");
    if (p->IF_InnerFlow != 0 && p->IF_InnerFlow != 1) return fail(OCFD_EARG, "IF_InnerFlow must be 0 or 1");
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (ndev < 1) return fail(OCFD_EHIP, "no HIP device");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    int r = scratch_setup(dev);
    if (r) return r;
    ocfd_ctx* c = new ocfd_ctx();
    c->hp = *p;
    c->p = to_par(*p);
    c->device = dev;
    HIPCHK(hipStreamCreateWithFlags(&c->e.st, hipStreamNonBlocking));
    *out = c;
    return OCFD_OK;
}

int ocfd_destroy(ocfd_ctx* c) {
    if (!c) return OCFD_OK;
    // a context destroyed while in a batch leaves it first: the batch's tables and graphs hold
    // its block arrays
    if (c->batch) {
        ocfd_batch* b = c->batch;
        (void)hipStreamSynchronize(b->e.st);
        pool_release(b, c);
        for (size_t q = 0; q < b->members.size(); q++)
            if (b->members[q] == c) { b->members.erase(b->members.begin() + q); break; }
        drop_plans(b->e);
        c->batch = nullptr;
    }
    c->e.du_pending = false;      // its DU goes with it
    engine_destroy(c->e);
    for (HostBlock& hb : c->blocks) for (void* q : hb.allocs) (void)hipFree(q);
    if (c->norm_buf) (void)hipFree(c->norm_buf);
    if (c->force_buf) (void)hipFree(c->force_buf);
    delete c;
    return OCFD_OK;
}

static int add_block_common(ocfd_ctx* c, int32_t no, int32_t nx, int32_t ny, int32_t nz, int32_t nsub, const int32_t* bc21, HostBlock** out) {
    if (!c || !bc21) return fail(OCFD_EARG, "null argument");
    if (no < 1 || nx < 3 || ny < 3 || nz < 3 || nsub < 1) return fail(OCFD_EARG, "block number or size out of range (need nx,ny,nz >= 3)");
    if ((int)c->blocks.size() < no) c->blocks.resize(no);
    HostBlock& hb = c->blocks[no - 1];
    if (hb.no != 0) return fail(OCFD_EARG, "block " + std::to_string(no) + " added twice");
    hb.no = no; hb.nx = nx; hb.ny = ny; hb.nz = nz;
    read_subs(hb, nsub, bc21);
    int r = alloc_block(c, hb);
    if (r) return r;
    c->bytes += hb.bytes;
    drop_plans(c->e);
    if (c->norm_buf) { (void)hipFree(c->norm_buf); c->norm_buf = nullptr; c->norm_bytes = 0; }
    *out = &hb;
    return OCFD_OK;
}

int ocfd_add_block(ocfd_ctx* c, int32_t no, int32_t nx, int32_t ny, int32_t nz, int32_t nsub, const int32_t* bc21,
                   const ocfd_block_geom* g) {
    if (!g) return fail(OCFD_EARG, "null geometry");
    HostBlock* hb;
    int r = add_block_common(c, no, nx, ny, nz, nsub, bc21, &hb);
    if (r) return r;
    Blk& b = hb->d;
    const int L = LAP;
    int n0[3] = {0, 0, 0}, n1[3] = {nx + 1, ny + 1, nz + 1};
    int c0[3] = {1 - L, 1 - L, 1 - L}, c1[3] = {nx + L - 1, ny + L - 1, nz + L - 1};
    int v0[3] = {1, 1, 1}, v1[3] = {nx - 1, ny - 1, nz - 1};
    int f1[3] = {nx, ny, nz};
    if (!g->x || !g->y || !g->z || !g->xc || !g->yc || !g->zc || !g->vol || !g->dtime_mesh || !g->Si || !g->Sj || !g->Sk)
        return fail(OCFD_EARG, "geometry pointer missing");
    if ((r = put_box(*hb, b.x, g->x, n0, n1)) || (r = put_box(*hb, b.y, g->y, n0, n1)) || (r = put_box(*hb, b.z, g->z, n0, n1))) return r;
    if ((r = put_box(*hb, b.xc, g->xc, c0, c1)) || (r = put_box(*hb, b.yc, g->yc, c0, c1)) || (r = put_box(*hb, b.zc, g->zc, c0, c1))) return r;
    if ((r = put_box(*hb, b.vol, g->vol, v0, v1)) || (r = put_box(*hb, b.dtm, g->dtime_mesh, v0, v1))) return r;
    if ((r = put_box(*hb, b.S[0], g->Si, v0, f1)) || (r = put_box(*hb, b.S[1], g->Sj, v0, f1)) || (r = put_box(*hb, b.S[2], g->Sk, v0, f1))) return r;
    for (int q = 0; q < 9; q++) {
        if (!g->n[q]) return fail(OCFD_EARG, "normal pointer missing");
        if ((r = put_box(*hb, b.nv[q], g->n[q], v0, f1))) return r;
    }
    for (int s = 1; s < 4; s++) for (int q = 0; q < 9; q++) {
        if (!g->met[s][q]) return fail(OCFD_EARG, "metric pointer missing");
        if ((r = put_box(*hb, b.met[s][q], g->met[s][q], v0, f1))) return r;
    }
    if (c->p.turb != 0) {
        if (!g->dw) return fail(OCFD_EARG, "wall distance required by the turbulence model");
        if ((r = put_box(*hb, b.dw, g->dw, v0, v1))) return r;
        for (int q = 0; q < 9; q++) {
            if (!g->met[0][q]) return fail(OCFD_EARG, "metric set 0 required by the turbulence model");
            if ((r = put_box(*hb, b.met[0][q], g->met[0][q], v0, f1))) return r;
        }
    }
    return upload_bcflags(*hb);
}

int ocfd_add_block_nodes(ocfd_ctx* c, int32_t no, int32_t nx, int32_t ny, int32_t nz, int32_t nsub, const int32_t* bc21,
                         const double* x, const double* y, const double* z) {
    if (!x || !y || !z) return fail(OCFD_EARG, "null coordinates");
    HostBlock* hb;
    int r = add_block_common(c, no, nx, ny, nz, nsub, bc21, &hb);
    if (r) return r;
    size_t nn = (size_t)nx * ny * nz;
    const double* src[3] = {x, y, z};
    for (int q = 0; q < 3; q++) hb->nodes[q].assign(src[q], src[q] + nn);
    return upload_bcflags(*hb);
}

// ghost node layer: Coordinate_send_mpi, Coordinate_Periodic, coordinate_boundary_and_corner and
// get_xyz_conner (sub_update_buffer_mpi.f90:250-512) on host copies of the node boxes
static void coordinate_ghosts(ocfd_ctx* c, std::vector<std::vector<double>> (&X)[3], double pd[3]) {
    for (size_t m = 0; m < c->blocks.size(); m++) {
        HostBlock& hb = c->blocks[m];
        for (const Sub& B : hb.subs) {
            if (B.bc >= 0) continue;
            int kb[3] = {B.ib, B.jb, B.kb}, ke[3] = {B.ie, B.je, B.ke};
            int kd = (B.face - 1) % 3;
            if (B.face <= 3) kb[kd] += 1; else kb[kd] -= 1;
            ke[kd] = kb[kd];
            int kb1[3] = {B.ib1, B.jb1, B.kb1}, ke1[3] = {B.ie1, B.je1, B.ke1};
            int kd1 = (B.face1 - 1) % 3;
            if (B.face1 <= 3) kb1[kd1] -= 1; else kb1[kd1] += 1;
            ke1[kd1] = kb1[kd1];
            int Ls[3] = {B.L1, B.L2, B.L3}, L[3], P[3], ks[3];
            for (int q = 0; q < 3; q++) { L[q] = std::abs(Ls[q]); P[q] = Ls[q] >= 0 ? 1 : -1; ks[q] = P[q] > 0 ? kb[q] : ke[q]; }
            HostBlock& t = c->blocks[B.nb1 - 1];
            for (int k = kb1[2]; k <= ke1[2]; k++) for (int j = kb1[1]; j <= ke1[1]; j++) for (int i = kb1[0]; i <= ke1[0]; i++) {
                int ka[3] = {i - kb1[0], j - kb1[1], k - kb1[2]};
                int i1 = ks[0] + ka[L[0] - 1] * P[0], j1 = ks[1] + ka[L[1] - 1] * P[1], k1 = ks[2] + ka[L[2] - 1] * P[2];
                for (int a = 0; a < 3; a++) X[a][B.nb1 - 1][IX(t.d, i, j, k)] = X[a][m][IX(hb.d, i1, j1, k1)];
            }
        }
    }
    for (size_t m = 0; m < c->blocks.size(); m++) {
        HostBlock& hb = c->blocks[m];
        for (const Sub& B : hb.subs) {
            if (B.bc != -2 && B.bc != -3) continue;
            double sgn = B.bc == -2 ? -1.0 : 1.0;
            int kb[3] = {B.ib, B.jb, B.kb}, ke[3] = {B.ie, B.je, B.ke};
            int kd = (B.face - 1) % 3;
            if (B.face <= 3) kb[kd] -= 1; else kb[kd] += 1;
            ke[kd] = kb[kd];
            for (int k = kb[2]; k <= ke[2]; k++) for (int j = kb[1]; j <= ke[1]; j++) for (int i = kb[0]; i <= ke[0]; i++)
                for (int a = 0; a < 3; a++) { double& v = X[a][m][IX(hb.d, i, j, k)]; v = v + (sgn < 0 ? -pd[a] : pd[a]); }
        }
    }
    for (size_t m = 0; m < c->blocks.size(); m++) {
        HostBlock& hb = c->blocks[m];
        const Blk& b = hb.d;
        for (const Sub& B : hb.subs) {
            if (B.bc < 0) continue;
            for (int a = 0; a < 3; a++) {
                std::vector<double>& x = X[a][m];
                for (int k = B.kb; k <= B.ke; k++) for (int j = B.jb; j <= B.je; j++) for (int i = B.ib; i <= B.ie; i++) {
                    switch (B.face) {
                    case 1: if (i == B.ib) x[IX(b, i - 1, j, k)] = 2.0 * x[IX(b, i, j, k)] - x[IX(b, i + 1, j, k)]; break;
                    case 2: if (j == B.jb) x[IX(b, i, j - 1, k)] = 2.0 * x[IX(b, i, j, k)] - x[IX(b, i, j + 1, k)]; break;
                    case 3: if (k == B.kb) x[IX(b, i, j, k - 1)] = 2.0 * x[IX(b, i, j, k)] - x[IX(b, i, j, k + 1)]; break;
                    case 4: if (i == B.ie) x[IX(b, i + 1, j, k)] = 2.0 * x[IX(b, i, j, k)] - x[IX(b, i - 1, j, k)]; break;
                    case 5: if (j == B.je) x[IX(b, i, j + 1, k)] = 2.0 * x[IX(b, i, j, k)] - x[IX(b, i, j - 1, k)]; break;
                    default: if (k == B.ke) x[IX(b, i, j, k + 1)] = 2.0 * x[IX(b, i, j, k)] - x[IX(b, i, j, k - 1)]; break;
                    }
                }
            }
        }
        int nx = b.nx, ny = b.ny, nz = b.nz;
        for (int a = 0; a < 3; a++) {
            std::vector<double>& x = X[a][m];
            auto V = [&](int i, int j, int k) -> double& { return x[IX(b, i, j, k)]; };
            for (int k = 1; k <= nz; k++) {
                V(0, 0, k) = (V(1, 0, k) + V(0, 1, k)) * 0.5; V(nx + 1, 0, k) = (V(nx, 0, k) + V(nx + 1, 1, k)) * 0.5;
                V(0, ny + 1, k) = (V(1, ny + 1, k) + V(0, ny, k)) * 0.5; V(nx + 1, ny + 1, k) = (V(nx, ny + 1, k) + V(nx + 1, ny, k)) * 0.5;
            }
            for (int j = 1; j <= ny; j++) {
                V(0, j, 0) = (V(1, j, 0) + V(0, j, 1)) * 0.5; V(nx + 1, j, 0) = (V(nx, j, 0) + V(nx + 1, j, 1)) * 0.5;
                V(0, j, nz + 1) = (V(1, j, nz + 1) + V(0, j, nz)) * 0.5; V(nx + 1, j, nz + 1) = (V(nx, j, nz + 1) + V(nx + 1, j, nz)) * 0.5;
            }
            for (int i = 1; i <= nx; i++) {
                V(i, 0, 0) = (V(i, 1, 0) + V(i, 0, 1)) * 0.5; V(i, ny + 1, 0) = (V(i, ny, 0) + V(i, ny + 1, 1)) * 0.5;
                V(i, 0, nz + 1) = (V(i, 1, nz + 1) + V(i, 0, nz)) * 0.5; V(i, ny + 1, nz + 1) = (V(i, ny, nz + 1) + V(i, ny + 1, nz)) * 0.5;
            }
            for (int cidx = 0; cidx < 8; cidx++) {
                int I = (cidx & 1) ? nx + 1 : 0, J = (cidx & 2) ? ny + 1 : 0, K = (cidx & 4) ? nz + 1 : 0;
                int Ii = (cidx & 1) ? nx : 1, Jj = (cidx & 2) ? ny : 1, Kk = (cidx & 4) ? nz : 1;
                V(I, J, K) = (2.0 * ((V(Ii, J, K) + V(I, Jj, K)) + V(I, J, Kk)) - ((V(Ii, Jj, K) + V(Ii, J, Kk)) + V(I, Jj, Kk))) / 3.0;
            }
        }
    }
}

int ocfd_setup_geometry(ocfd_ctx* c) {
    int r = ready(c);
    if (r) return r;
    for (HostBlock& hb : c->blocks) {
        if (hb.nodes[0].empty()) return fail(OCFD_EARG, "ocfd_setup_geometry needs every block added with ocfd_add_block_nodes");
        if ((r = check_subs(c, hb))) return r;
    }
    for (HostBlock& hb : c->blocks) hb.geof_ok = false;
    std::vector<std::vector<double>> X[3];
    for (int a = 0; a < 3; a++) {
        X[a].resize(c->blocks.size());
        for (size_t m = 0; m < c->blocks.size(); m++) {
            HostBlock& hb = c->blocks[m];
            X[a][m].assign(hb.n(), 0.0);
            for (int k = 1; k <= hb.nz; k++) for (int j = 1; j <= hb.ny; j++) for (int i = 1; i <= hb.nx; i++)
                X[a][m][IX(hb.d, i, j, k)] = hb.nodes[a][(size_t)(i - 1) + (size_t)hb.nx * ((size_t)(j - 1) + (size_t)hb.ny * (k - 1))];
        }
    }
    double pd[3] = {c->hp.Periodic_d[0], c->hp.Periodic_d[1], c->hp.Periodic_d[2]};
    coordinate_ghosts(c, X, pd);
    for (size_t m = 0; m < c->blocks.size(); m++) {
        HostBlock& hb = c->blocks[m];
        HIPCHK(hipMemcpy(hb.d.x, X[0][m].data(), hb.n() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(hb.d.y, X[1][m].data(), hb.n() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(hb.d.z, X[2][m].data(), hb.n() * sizeof(double), hipMemcpyHostToDevice));
        launch_geometry(hb.d, hb.subs.data(), (int)hb.subs.size(), c->e.st);
    }
    // update_Mesh_Center: cell-centre ghosts of inner faces, then the periodic shift
    // (sub_udate_Meshlink_mpi.f90:12-249), through the U-halo kernel on (xc, yc, zc)
    for (HostBlock& hb : c->blocks)
        for (const Sub& B : hb.subs) {
            if (B.bc >= 0) continue;
            Blk s = hb.d, t = c->blocks[B.nb1 - 1].d;
            s.U[0] = hb.d.xc; s.U[1] = hb.d.yc; s.U[2] = hb.d.zc;
            t.U[0] = t.xc; t.U[1] = t.yc; t.U[2] = t.zc;
            launch_halo(s, t, B, 3, c->e.st);
        }
    HIPCHK(hipStreamSynchronize(c->e.st));
    for (HostBlock& hb : c->blocks) {
        bool any = false;
        for (const Sub& B : hb.subs) any = any || B.bc == -2 || B.bc == -3;
        if (!any) continue;
        std::vector<double> C[3];
        double* dev[3] = {hb.d.xc, hb.d.yc, hb.d.zc};
        for (int a = 0; a < 3; a++) if ((r = get_box(hb, dev[a], C[a]))) return r;
        for (const Sub& B : hb.subs) {
            if (B.bc != -2 && B.bc != -3) continue;
            int kb[3] = {B.ib, B.jb, B.kb}, ke[3] = {B.ie - 1, B.je - 1, B.ke - 1};
            int kd = (B.face - 1) % 3;
            if (B.face <= 3) kb[kd] -= LAP;
            ke[kd] = kb[kd] + LAP - 1;
            for (int k = kb[2]; k <= ke[2]; k++) for (int j = kb[1]; j <= ke[1]; j++) for (int i = kb[0]; i <= ke[0]; i++)
                for (int a = 0; a < 3; a++) { double& v = C[a][IX(hb.d, i, j, k)]; v = v + (B.bc == -2 ? -pd[a] : pd[a]); }
        }
        for (int a = 0; a < 3; a++) HIPCHK(hipMemcpy(dev[a], C[a].data(), hb.n() * sizeof(double), hipMemcpyHostToDevice));
    }
    // wall distance (sub_comput_dw.f90:43-168): wall nodes of every block, in block/subface order
    if (c->p.turb != 0) {
        std::vector<double> W;
        for (size_t m = 0; m < c->blocks.size(); m++) {
            HostBlock& hb = c->blocks[m];
            for (const Sub& B : hb.subs) {
                if (B.bc != 2) continue;
                for (int k = B.kb; k <= B.ke; k++) for (int j = B.jb; j <= B.je; j++) for (int i = B.ib; i <= B.ie; i++)
                    for (int a = 0; a < 3; a++) W.push_back(X[a][m][IX(hb.d, i, j, k)]);
            }
        }
        double* dW = nullptr;
        HIPCHK(hipMalloc(&dW, std::max<size_t>(W.size(), 3) * sizeof(double)));
        if (!W.empty()) HIPCHK(hipMemcpy(dW, W.data(), W.size() * sizeof(double), hipMemcpyHostToDevice));
        for (HostBlock& hb : c->blocks) launch_wall_distance(hb.d, dW, (long)(W.size() / 3), c->e.st);
        HIPCHK(hipStreamSynchronize(c->e.st));
        HIPCHK(hipFree(dW));
    }
    for (HostBlock& hb : c->blocks) launch_mesh_quality(hb.d, c->e.st);
    for (HostBlock& hb : c->blocks) for (int q = 0; q < 3; q++) hb.nodes[q].clear();
    drop_plans(c->e);        // the sweep-ordered geometry is taken from the new arrays
    return finish(c);
}

int ocfd_upload_U(ocfd_ctx* c, int32_t no, const double* U) {
    HostBlock* hb = find_block(c, no);
    if (!hb || !U) return fail(OCFD_EARG, "unknown block or null U");
    int NV = c->p.NV;
    const Blk& b = hb->d;
    std::vector<double> h(hb->n());
    long ni = b.NI, nj = b.NJ;
    for (int m = 0; m < NV; m++) {
        for (long s = 0; s < (long)hb->n(); s++) h[s] = U[m + (long)NV * s];   // U(m,i,j,k) over the same LAP box
        HIPCHK(hipMemcpy(b.U[m], h.data(), hb->n() * sizeof(double), hipMemcpyHostToDevice));
    }
    (void)ni; (void)nj;
    return OCFD_OK;
}

int ocfd_download_U(ocfd_ctx* c, int32_t no, double* U) {
    HostBlock* hb = find_block(c, no);
    if (!hb || !U) return fail(OCFD_EARG, "unknown block or null U");
    HIPCHK(hipStreamSynchronize(c->e.st));
    int NV = c->p.NV, r;
    std::vector<double> h;
    for (int m = 0; m < NV; m++) {
        if ((r = get_box(*hb, hb->d.U[m], h))) return r;
        for (long s = 0; s < (long)hb->n(); s++) U[m + (long)NV * s] = h[s];
    }
    return OCFD_OK;
}

int ocfd_download_field(ocfd_ctx* c, int32_t no, int32_t field, double* out) {
    HostBlock* hb = find_block(c, no);
    if (!hb || !out) return fail(OCFD_EARG, "unknown block or null output");
    HIPCHK(hipStreamSynchronize(c->e.st));
    if (field == OCFD_F_DU) {
        flush_du(c->batch ? c->batch->e : c->e);
        HIPCHK(hipGetLastError());
    }
    const Blk& b = hb->d;
    int nv = (field == OCFD_F_RES || field == OCFD_F_DU) ? c->p.NV : 1;
    long n0 = b.nx + 3, n1 = b.ny + 3, n2 = b.nz + 3;
    std::vector<double> h;
    int r;
    for (int m = 0; m < nv; m++) {
        const double* src;
        switch (field) {
        case OCFD_F_RES: src = b.Res[m]; break;
        case OCFD_F_DU: src = b.DU[m]; break;
        case OCFD_F_DT: src = b.dt; break;
        case OCFD_F_MU: src = b.muf; break;
        case OCFD_F_MUT: src = b.mut; break;
        case OCFD_F_DW:
            if (!b.dw) return fail(OCFD_EARG, "no wall distance without a turbulence model");
            src = b.dw; break;
        default: return fail(OCFD_EARG, "unknown field id");
        }
        if ((r = get_box(*hb, src, h))) return r;
        for (int k = -1; k <= b.nz + 1; k++) for (int j = -1; j <= b.ny + 1; j++) for (int i = -1; i <= b.nx + 1; i++)
            out[m + nv * ((i + 1) + n0 * ((j + 1) + n1 * (long)(k + 1)))] = h[IX(b, i, j, k)];
    }
    (void)n2;
    return OCFD_OK;
}

int ocfd_init_flow(ocfd_ctx* c) {
    int r = ready_solo(c);
    if (r) return r;
    // init_flow_zero's external-flow branch; the inner-flow initial field is outside this path
    if (c->hp.IF_InnerFlow == 1) return fail(OCFD_EUNSUPPORTED, "ocfd_init_flow: the inner-flow initial field (IF_InnerFlow = 1) is outside this path");
    double g = c->p.gamma, Ma = c->p.Ma;
    double d0 = 1.0, p0 = 1.0 / (g * Ma * Ma);
    double u0 = std::cos(c->p.A_alfa) * std::cos(c->p.A_beta), v0 = std::sin(c->p.A_alfa) * std::cos(c->p.A_beta), w0 = std::sin(c->p.A_beta);
    double q[7] = {d0, d0 * u0, d0 * v0, d0 * w0, p0 / (g - 1.0) + (0.5 * d0) * ((u0 * u0 + v0 * v0) + w0 * w0), 0.0, 0.0};
    if (c->p.NV == 6) q[5] = 5.0;
    if (c->p.NV == 7) { q[5] = 10.0 * c->p.Kt_inf; q[6] = c->p.Wt_inf; }
    for (HostBlock& hb : c->blocks) {
        if ((r = check_subs(c, hb))) return r;
        std::vector<double> h(hb.n());
        for (int m = 0; m < c->p.NV; m++) {
            std::fill(h.begin(), h.end(), q[m]);
            HIPCHK(hipMemcpy(hb.d.U[m], h.data(), hb.n() * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    enqueue_bc(c);
    enqueue_halo(c);
    return finish(c);
}

int ocfd_step_lusgs(ocfd_ctx* c, int32_t niter) {
    int r = prepare(c);
    if (r) return r;
    for (HostBlock& hb : c->blocks) if ((r = check_subs(c, hb))) return r;
    if ((r = engine_step(c->e, c->refs(), c->p, niter, nullptr))) return r;
    for (int it = 0; it < niter; it++) { c->tt = c->tt + c->p.dt_global; c->kstep = c->kstep + 1; }
    return OCFD_OK;
}

int ocfd_step_timed(ocfd_ctx* c, int32_t niter, double* stage_ms) {
    int r = prepare(c);
    if (r) return r;
    if (!stage_ms || niter < 1) return fail(OCFD_EARG, "null stage_ms or niter < 1");
    for (HostBlock& hb : c->blocks) if ((r = check_subs(c, hb))) return r;
    if ((r = engine_step(c->e, c->refs(), c->p, niter, stage_ms))) return r;
    for (int it = 0; it < niter; it++) { c->tt = c->tt + c->p.dt_global; c->kstep = c->kstep + 1; }
    return OCFD_OK;
}

int ocfd_residual(ocfd_ctx* c) { int r = prepare(c); if (r) return r; enqueue_residual(c); return finish(c); }
int ocfd_update_limit(ocfd_ctx* c) { int r = ready_solo(c); if (r) return r; enqueue_stage(c, OCFD_STAGE_UPDATE); return finish(c); }
int ocfd_bc(ocfd_ctx* c) {
    int r = ready_solo(c); if (r) return r;
    for (HostBlock& hb : c->blocks) if ((r = check_subs(c, hb))) return r;
    enqueue_bc(c); return finish(c);
}
int ocfd_halo(ocfd_ctx* c) { int r = ready_solo(c); if (r) return r; enqueue_halo(c); return finish(c); }
int ocfd_force_vt_kw(ocfd_ctx* c) { int r = ready_solo(c); if (r) return r; enqueue_force_vt_kw(c); return finish(c); }

int ocfd_res_norms(ocfd_ctx* c, double* res_max, double* res_rms) {
    int r = ready(c);
    if (r) return r;
    if (!res_max || !res_rms) return fail(OCFD_EARG, "null output");
    const int nb = (int)c->blocks.size();
    const NormLayout L = norm_layout(nb, c->p.NV);
    if (c->norm_bytes < L.bytes) {
        if (c->norm_buf) HIPCHK(hipFree(c->norm_buf));
        c->norm_buf = nullptr; c->norm_bytes = 0;
        HIPCHK(hipMalloc((void**)&c->norm_buf, L.bytes));
        c->norm_bytes = L.bytes;
        std::vector<NormBlk> tab((size_t)nb);
        fill_norm_table(c, tab.data());
        HIPCHK(hipMemcpyAsync(c->norm_buf + L.tab, tab.data(), (size_t)nb * sizeof(NormBlk), hipMemcpyHostToDevice, c->e.st));
    }
    return run_norms(c, c->norm_buf, L, res_max, res_rms);
}

int ocfd_res_norms_noalloc(ocfd_ctx* c, double* res_max, double* res_rms) {
    if (!c) return fail(OCFD_EARG, "null context");
    char* dev = scratch_of(c->device);
    if (!dev) return fail(OCFD_EHIP, "no scratch address for the context's device");
    return norms_into(c, dev, SCRATCH_BYTES, res_max, res_rms);
}

int ocfd_forces(ocfd_ctx* c, double* out) {
    int r = ready(c);
    if (r) return r;
    if (!out) return fail(OCFD_EARG, "null output");
    const SetPlan& P = c->e.set;
    const size_t need = ((size_t)P.force.ntiles + 1) * 9;
    if (c->force_n < need) {
        if (c->force_buf) HIPCHK(hipFree(c->force_buf));
        c->force_buf = nullptr; c->force_n = 0;
        HIPCHK(hipMalloc((void**)&c->force_buf, need * sizeof(double)));
        c->force_n = need;
    }
    const double cen[3] = {c->hp.Centroid[0], c->hp.Centroid[1], c->hp.Centroid[2]};
    double* out9 = c->force_buf + (size_t)P.force.ntiles * 9;
    launch_forces(P, cen, c->force_buf, out9, c->e.st);
    HIPCHK(hipGetLastError());
    std::vector<double> tot(9, 0.0);
    HIPCHK(hipMemcpyAsync(tot.data(), out9, 9 * sizeof(double), hipMemcpyDeviceToHost, c->e.st));
    HIPCHK(hipStreamSynchronize(c->e.st));
    for (int q = 0; q < 6; q++) tot[q] = 2.0 * tot[q] / c->hp.Ref_S;
    for (int q = 6; q < 9; q++) tot[q] = 2.0 * tot[q] / (c->hp.Ref_S * c->hp.Ref_L);
    double Fx = tot[0] + tot[3], Fy = tot[1] + tot[4], Fz = tot[2] + tot[5], AoA = c->hp.AoA;
    for (int q = 0; q < 9; q++) out[q] = tot[q];
    if (c->hp.Cood_Y_UP == 1) { out[9] = Fy * std::cos(AoA) - Fx * std::sin(AoA); out[10] = Fx * std::cos(AoA) + Fy * std::sin(AoA); out[11] = Fz; }
    else { out[9] = Fz * std::cos(AoA) - Fx * std::sin(AoA); out[10] = Fx * std::cos(AoA) + Fz * std::sin(AoA); out[11] = Fy; }
    return OCFD_OK;
}

int ocfd_block_status(ocfd_ctx* c, int32_t no, int32_t* overlimit, int32_t* nlimited) {
    HostBlock* hb = find_block(c, no);
    if (!hb || !overlimit || !nlimited) return fail(OCFD_EARG, "unknown block or null output");
    int f[16];
    HIPCHK(hipStreamSynchronize(c->e.st));
    HIPCHK(hipMemcpy(f, hb->d.flags, sizeof(f), hipMemcpyDeviceToHost));
    *overlimit = f[0]; *nlimited = f[1];
    return OCFD_OK;
}

int ocfd_get_counters(ocfd_ctx* c, int32_t* kstep, double* tt) {
    if (!c || !kstep || !tt) return fail(OCFD_EARG, "null argument");
    *kstep = c->kstep; *tt = c->tt;
    return OCFD_OK;
}

int ocfd_set_graph(ocfd_ctx* c, int32_t enable) {
    if (!c) return fail(OCFD_EARG, "null context");
    if ((enable != 0) != c->e.use_graph) drop_graphs(c->e);
    c->e.use_graph = enable != 0;
    return OCFD_OK;
}

int ocfd_set_lu_schedule(ocfd_ctx* c, int32_t schedule) {
    if (!c) return fail(OCFD_EARG, "null context");
    return engine_set_lu_schedule(c->e, schedule);
}

int ocfd_set_lu_fail_waits(ocfd_ctx* c, int32_t on) {
    if (!c) return fail(OCFD_EARG, "null context");
    Engine& e = c->e;
    e.lu_fail = on ? 1 : 0;
    // the sweeps read the word at run time: captured graphs stay valid
    HIPCHK(hipStreamSynchronize(e.st));
    if (e.lu.err) HIPCHK(hipMemcpy(e.lu.err + 2, &e.lu_fail, sizeof(int), hipMemcpyHostToDevice));
    for (Engine& q : e.parts) {
        q.lu_fail = e.lu_fail;
        if (q.lu.err) HIPCHK(hipMemcpy(q.lu.err + 2, &q.lu_fail, sizeof(int), hipMemcpyHostToDevice));
    }
    return OCFD_OK;
}

int ocfd_set_per_block(ocfd_ctx* c, int32_t on) {
    if (!c) return fail(OCFD_EARG, "null context");
    Engine& e = c->e;
    if ((on != 0) == e.per_block) return OCFD_OK;
    if (on && e.lu_tj != 8) { int r = engine_set_lu_tile(e, 8); if (r) return r; }
    HIPCHK(hipStreamSynchronize(e.st));
    flush_du(e);
    drop_graphs(e);
    drop_parts(e);
    e.per_block = on != 0;
    return OCFD_OK;
}

int ocfd_set_lu_tile(ocfd_ctx* c, int32_t tj) {
    if (!c) return fail(OCFD_EARG, "null context");
    return engine_set_lu_tile(c->e, tj);
}

int ocfd_get_lu_tile(ocfd_ctx* c, int32_t* tj) {
    if (!c || !tj) return fail(OCFD_EARG, "null context or output");
    *tj = c->e.lu_tj;
    return OCFD_OK;
}

int ocfd_autotune_lu(ocfd_ctx* c, int32_t niter, int32_t* best_tj, double* ms) {
    int r = prepare(c);
    if (r) return r;
    for (HostBlock& hb : c->blocks) if ((r = check_subs(c, hb))) return r;
    int best = 0;
    if ((r = engine_autotune(c->e, c->refs(), c->p, niter, &best, ms))) return r;
    if (best_tj) *best_tj = best;
    for (int it = 0; it < 3 * niter; it++) { c->tt = c->tt + c->p.dt_global; c->kstep = c->kstep + 1; }
    return OCFD_OK;
}

int ocfd_sync(ocfd_ctx* c) {
    if (!c) return fail(OCFD_EARG, "null context");
    HIPCHK(hipStreamSynchronize(c->e.st));
    return OCFD_OK;
}

// ---- batches -------------------------------------------------------------------------------

static bool same_switches(const Par& a, const Par& b) {
    return a.NV == b.NV && a.visc == b.visc && a.turb == b.turb && a.scheme == b.scheme && a.flux == b.flux &&
           a.recon == b.recon && a.bscheme == b.bscheme && a.local_dt == b.local_dt && a.dtime_mesh == b.dtime_mesh &&
           a.positivity == b.positivity && a.limit_flow == b.limit_flow && a.gamma == b.gamma && a.w_LU == b.w_LU &&
           a.mixed == b.mixed;
}

int ocfd_batch_create(ocfd_batch** out) {
    if (!out) return fail(OCFD_EARG, "null argument");
    ocfd_batch* b = new ocfd_batch();
    b->e.owner = b;
    hipError_t e = hipStreamCreateWithFlags(&b->e.st, hipStreamNonBlocking);
    if (e != hipSuccess) { delete b; return fail(OCFD_EHIP, std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e)); }
    *out = b;
    return OCFD_OK;
}

int ocfd_batch_destroy(ocfd_batch* b) {
    if (!b) return OCFD_OK;
    engine_destroy(b->e);
    pool_destroy(b);
    for (ocfd_ctx* c : b->members) c->batch = nullptr;     // the members may be stepped alone again
    delete b;
    return OCFD_OK;
}

int ocfd_batch_add(ocfd_batch* b, ocfd_ctx* c) {
    if (!b || !c) return fail(OCFD_EARG, "null argument");
    int r = ready(c);
    if (r) return r;
    for (HostBlock& hb : c->blocks) if ((r = check_subs(c, hb))) return r;
    if (c->batch == b) return fail(OCFD_EARG, "context already in the batch");
    if (c->batch) return fail(OCFD_EARG, "context already in another batch");
    if (!b->members.empty() && !same_switches(b->members[0]->p, c->p))
        return fail(OCFD_EARG, "a batch needs the same NVAR, schemes, flux, turbulence model, limiter, gamma, w_LU and precision in every member");
    HIPCHK(hipStreamSynchronize(c->e.st));
    b->members.push_back(c);
    c->batch = b;
    drop_plans(b->e);
    // the member steps through the batch's launches now: its own tables (the sweep storage is
    // about 380 B per cell) are rebuilt only if it is stepped alone again
    drop_plans(c->e);
    return OCFD_OK;
}

int ocfd_batch_remove(ocfd_batch* b, ocfd_ctx* c) {
    if (!b || !c) return fail(OCFD_EARG, "null argument");
    for (size_t q = 0; q < b->members.size(); q++) {
        if (b->members[q] != c) continue;
        HIPCHK(hipStreamSynchronize(b->e.st));
        pool_release(b, c);
        b->members.erase(b->members.begin() + q);
        c->batch = nullptr;
        drop_plans(b->e);
        return OCFD_OK;
    }
    return fail(OCFD_EARG, "context not in the batch");
}

static int batch_step(ocfd_batch* b, int32_t niter, double* stage_ms) {
    if (!b) return fail(OCFD_EARG, "null batch");
    if (b->members.empty()) return OCFD_OK;
    if (stage_ms && niter < 1) return fail(OCFD_EARG, "niter < 1");
    const Par& p0 = b->members[0]->p;
    std::vector<BlockRef> refs = b->refs();
    int r = engine_prepare(b->e, refs, p0, true);
    if (r) return r;
    if ((r = engine_step(b->e, refs, p0, niter, stage_ms))) return r;
    for (ocfd_ctx* c : b->members)
        for (int it = 0; it < niter; it++) { c->tt = c->tt + c->p.dt_global; c->kstep = c->kstep + 1; }
    return OCFD_OK;
}

int ocfd_batch_step(ocfd_batch* b, int32_t niter) { return batch_step(b, niter, nullptr); }

int ocfd_batch_step_timed(ocfd_batch* b, int32_t niter, double* stage_ms) {
    if (!stage_ms) return fail(OCFD_EARG, "null stage_ms");
    return batch_step(b, niter, stage_ms);
}

int ocfd_batch_set_lu_tile(ocfd_batch* b, int32_t tj) {
    if (!b) return fail(OCFD_EARG, "null batch");
    return engine_set_lu_tile(b->e, tj);
}

int ocfd_batch_autotune_lu(ocfd_batch* b, int32_t niter, int32_t* best_tj, double* ms) {
    if (!b) return fail(OCFD_EARG, "null batch");
    if (b->members.empty()) return fail(OCFD_EARG, "empty batch");
    const Par& p0 = b->members[0]->p;
    std::vector<BlockRef> refs = b->refs();
    int best = 0, r;
    if ((r = engine_autotune(b->e, refs, p0, niter, &best, ms))) return r;
    if (best_tj) *best_tj = best;
    for (ocfd_ctx* c : b->members)
        for (int it = 0; it < 3 * niter; it++) { c->tt = c->tt + c->p.dt_global; c->kstep = c->kstep + 1; }
    return OCFD_OK;
}

int ocfd_batch_set_graph(ocfd_batch* b, int32_t enable) {
    if (!b) return fail(OCFD_EARG, "null batch");
    if ((enable != 0) != b->e.use_graph) drop_graphs(b->e);
    b->e.use_graph = enable != 0;
    return OCFD_OK;
}

int ocfd_check_division(int64_t n, uint64_t seed, int32_t numerator_exponent, int64_t* mismatches) {
    if (n < 0 || !mismatches || numerator_exponent < 0 || numerator_exponent > 1000) return fail(OCFD_EARG, "bad argument");
    long long m = 0;
    if (check_division((long long)n, (unsigned long long)seed, numerator_exponent, &m)) return fail(OCFD_EHIP, "division check failed to run");
    *mismatches = (int64_t)m;
    return OCFD_OK;
}

int ocfd_mem_bytes(ocfd_ctx* c, int64_t* bytes) {
    if (!c || !bytes) return fail(OCFD_EARG, "null argument");
    *bytes = (int64_t)c->bytes;
    return OCFD_OK;
}

}  // extern "C"
