// Host launchers of the hot-path kernels (ocfd_kernels.hip), called by the runtime.
#pragma once
#include <vector>

#include "ocfd_dev.h"

namespace ocfd {
void launch_reset_flags(const Blk& b, hipStream_t st);
void launch_prims(const Blk& b, const Par& p, hipStream_t st);
void launch_viscosity(const Blk& b, const Par& p, hipStream_t st);
void launch_sa(const Blk& b, const Par& p, const Sub* subs, int nsub, hipStream_t st);
void launch_sst(const Blk& b, const Par& p, const Sub* subs, int nsub, hipStream_t st);
void launch_laminar_amut(const Blk& b, const Sub* subs, int nsub, hipStream_t st);
void launch_inv_fluxes(const Blk& b, const Par& p, const Sub* subs, int nsub, hipStream_t st);
void launch_visc_assemble_dt(const Blk& b, const Par& p, hipStream_t st);
void launch_lusgs_sweep(const Blk& b, const Par& p, bool forward, hipStream_t st);
std::vector<LuTask> lusgs_tasks(const std::vector<Blk>& blks, bool forward, long long* total_elements);
int lusgs_fields(int NV);
int lusgs_windows(int max_cells_i);
void launch_lusgs_pack_geometry(const LuPlan& L, int NV, hipStream_t st);
void launch_lusgs_pipelined(const LuPlan& L, const Par& p, bool forward, bool fenced, hipStream_t st);
void launch_update_limit(const Blk& b, const Par& p, hipStream_t st);
void launch_bc(const Blk& b, const Par& p, const Sub& B, hipStream_t st);
void launch_halo(const Blk& src, const Blk& dst, const Sub& B, int NV, hipStream_t st);
void launch_corners(const Blk& b, int NV, hipStream_t st);
void launch_force_vt_kw(const Blk& b, const Par& p, hipStream_t st);
int norms_parts();
void launch_norms(const Blk& b, int NV, double* part, hipStream_t st);
void launch_force_sub(const Blk& b, const Par& p, const Sub& B, double* acc12, hipStream_t st);
// device geometry precompute (ocfd_geometry.hip)
void launch_geometry(const Blk& b, const Sub* subs, int nsub, hipStream_t st);
void launch_mesh_quality(const Blk& b, hipStream_t st);
void launch_wall_distance(const Blk& b, const double* wall_xyz, long npw, hipStream_t st);
}  // namespace ocfd
