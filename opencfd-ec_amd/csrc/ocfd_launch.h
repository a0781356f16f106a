// Host launchers of the hot-path kernels (ocfd_kernels.hip), called by the runtime.
#pragma once
#include <vector>

#include "ocfd_dev.h"
#include "../../include/ocfd_ec.h"

namespace ocfd {

// task list of a subface launch: entries and, per tile, the entry it belongs to
struct SubList {
    SubTask* t = nullptr;
    int* tile_task = nullptr;
    int ntiles = 0;
};
struct HaloList {
    HaloTask* t = nullptr;
    int* tile_task = nullptr;
    int ntiles = 0;
};

// device tables of one block set (built by the runtime from its blocks)
struct SetPlan {
    Blk* blk = nullptr;
    Par* par = nullptr;
    int nb = 0;
    Par p0{};                        // the set-wide switches (turbulence model, NVAR, limiter)
    int* tile_blk[SP_COUNT] = {};
    int* tile0[SP_COUNT] = {};
    int ntiles[SP_COUNT] = {};
    SubList bc, amut, kbc, force;   // force: the wall (bc 2) subfaces, one thread per face
    SubList vghost;                  // the six faces of every block (k_visc_ghost)
    HaloList halo;
    int mx = 0;                      // max over blocks of max(nx, ny, nz)
    const int* halt = nullptr;       // the LU-SGS error word of the engine's sweeps (or null)
};

void set_pre(const SetPlan& P, hipStream_t st);
void set_inviscid(const SetPlan& P, hipStream_t st);
void set_viscous_dt(const SetPlan& P, hipStream_t st);
void set_update_limit(const SetPlan& P, hipStream_t st);
// the update taking DU from the pipelined sweeps' storage
void set_update_limit_sweep_order(const SetPlan& P, const LuPlan& L, hipStream_t st);
void set_bc(const SetPlan& P, hipStream_t st);
void set_halo(const SetPlan& P, hipStream_t st);
void set_force_vt_kw(const SetPlan& P, hipStream_t st);
// the residual norms of a table of blocks: parts (nb x NV x norms_parts() x 3 doubles), then
// out (3 NV doubles: max, sum of squares, NaN flag per component)
constexpr int NORMS_PARTS = 64;
inline int norms_parts() { return NORMS_PARTS; }
void launch_norms(const NormBlk* tab, int nb, int NV, double* part, double* out, hipStream_t st);
// the residual norms of a context in a caller-provided device buffer (ocfd_runtime.hip)
constexpr int NORMS_INTO_MAX_BLOCKS = 64;
int norms_into(ocfd_ctx* c, char* dev, size_t cap, double* res_max, double* res_rms);
int ctx_device(const ocfd_ctx* c);
int set_error(int code, const char* msg);

HaloMap halo_map(const Sub& B);
int halo_points(const HaloMap& h);
void launch_halo(const Blk& src, const Blk& dst, const Sub& B, int NV, hipStream_t st);
void launch_lusgs_sweep(const Blk& b, const Par& p, bool forward, hipStream_t st);
// comput_force over the set's wall faces: parts (ntiles x 9 doubles), then out9 (Ft(1:9))
void launch_forces(const SetPlan& P, const double cen[3], double* part, double* out9, hipStream_t st);
int check_division(long long n, unsigned long long seed, int emax, long long* mismatches);
std::vector<LuTask> lusgs_tasks(const std::vector<Blk>& blks, bool forward, int nfields, int tj, long long* total_elements);
bool lusgs_tile_ok(int tj);
int lusgs_fields(int NV);
// (forward tile, first i) of every i-window of the tiles, tile after tile
std::vector<int2> lusgs_window_list(const std::vector<Blk>& blks, const std::vector<LuTask>& fwd);
void launch_lusgs_pack_geometry(const LuPlan& L, const Par& p, hipStream_t st);
// the sweeps' per-iteration inputs (U, Res, Lc, Lv, the diagonal) into sweep order
void launch_lusgs_pack(const LuPlan& L, const Par& p, hipStream_t st);
void launch_lusgs_reset(const LuPlan& L, int NV, hipStream_t st);
// the backward sweep's DU from the sweep-ordered storage into the block arrays (valid from the
// end of the backward sweep until the next pack)
void launch_lusgs_du_to_box(const LuPlan& L, int NV, hipStream_t st);
void launch_lusgs_pipelined(const LuPlan& L, const Par& p, bool forward, bool fenced, hipStream_t st);
// device geometry precompute (ocfd_geometry.hip)
void launch_geometry(const Blk& b, const Sub* subs, int nsub, hipStream_t st);
void launch_mesh_quality(const Blk& b, hipStream_t st);
// the FP32 copies of the geometry for the mixed-precision kernels (a block of a mixed context)
void launch_geometry_f32(const Blk& b, hipStream_t st);
void launch_wall_distance(const Blk& b, const double* wall_xyz, long npw, hipStream_t st);
}  // namespace ocfd
