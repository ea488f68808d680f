// C ABI of the standalone remesh interpolation (host pointers).
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "fs_remesh.h"

namespace fs {
namespace {

struct Staged {
    int32_t* rec;
    double *ox, *op, *of;
};

int stage_old_mesh(fs_ctx* ctx, const int32_t* panels, int64_t n_panels, const double* ox,
                   const double* op, const double* of, int64_t n_old, size_t extra_bytes,
                   Staged& st, char*& extra) {
    const size_t brec = align_up(sizeof(int32_t) * remesh::kPanelInts * n_panels, 256);
    const size_t bpt = align_up(sizeof(double) * n_old, 256);
    FS_TRY(ctx->io.ensure(brec + 3 * bpt + extra_bytes, ctx->stream));
    char* s = ctx->io.as<char>();
    st.rec = reinterpret_cast<int32_t*>(s);
    s += brec;
    st.ox = reinterpret_cast<double*>(s);
    s += bpt;
    st.op = reinterpret_cast<double*>(s);
    s += bpt;
    st.of = reinterpret_cast<double*>(s);
    s += bpt;
    extra = s;
    FS_HIP_CHECK(hipMemcpyAsync(st.rec, panels, sizeof(int32_t) * remesh::kPanelInts * n_panels,
                                hipMemcpyHostToDevice, ctx->stream));
    FS_HIP_CHECK(hipMemcpyAsync(st.ox, ox, sizeof(double) * n_old, hipMemcpyHostToDevice,
                                ctx->stream));
    FS_HIP_CHECK(hipMemcpyAsync(st.op, op, sizeof(double) * n_old, hipMemcpyHostToDevice,
                                ctx->stream));
    FS_HIP_CHECK(hipMemcpyAsync(st.of, of, sizeof(double) * n_old, hipMemcpyHostToDevice,
                                ctx->stream));
    return FS_OK;
}

// Checks that every point index of every panel lies in [0, n_old) and every
// neighbour/child index in [-2, n_panels): the device search trusts them.
int check_panels(const int32_t* rec, int64_t n_panels, int64_t n_old) {
    if (n_panels < 1 || n_old < 9) return fail(FS_E_INVALID, "remesh: empty old mesh");
    for (int64_t i = 0; i < n_panels; ++i) {
        const int32_t* r = rec + i * remesh::kPanelInts;
        for (int k = 1; k <= 9; ++k)
            if (r[k] < 0 || r[k] >= n_old)
                return fail(FS_E_INVALID, "remesh: panel %lld point %d out of range",
                            (long long)i, k);
        for (int k = 13; k <= 16; ++k)
            if (r[k] < -2 || r[k] >= n_panels)
                return fail(FS_E_INVALID, "remesh: panel %lld neighbour out of range",
                            (long long)i);
        if (r[11] < -1 || r[11] >= n_panels)
            return fail(FS_E_INVALID, "remesh: panel %lld parent out of range", (long long)i);
        if (r[17] & (remesh::kFlagRefinedXP | remesh::kFlagRefinedP)) {
            const int nch = (r[17] & remesh::kFlagRefinedXP) ? 4 : 2;
            if (r[18] < 1 || r[18] + nch > n_panels)
                return fail(FS_E_INVALID, "remesh: panel %lld children out of range",
                            (long long)i);
        }
    }
    return FS_OK;
}

int read_status(fs_ctx* ctx, const int* d_status) {
    int status = 0;
    FS_HIP_CHECK(hipMemcpyAsync(&status, d_status, sizeof(int), hipMemcpyDeviceToHost,
                                ctx->stream));
    FS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return remesh::status_to_error(status);
}

int run_job(fs_ctx* ctx, remesh::Job& j, int npanels, int n, int np, bool grid,
            double* f_host) {
    // job descriptor and its work arrays live in the context's scratch
    const size_t bmesh = remesh::mesh_bytes(npanels);
    const size_t bwork = remesh::grid_bytes(n, np);
    const size_t bjob = align_up(sizeof(remesh::Job), 256);
    FS_TRY(ctx->scratch.ensure(bmesh + bwork + bjob, ctx->stream));
    remesh::carve_mesh(j, ctx->scratch.ptr, npanels);
    remesh::carve_grid(j, ctx->scratch.as<char>(bmesh), n, np);
    remesh::Job* d_job = ctx->scratch.as<remesh::Job>(bmesh + bwork);
    FS_HIP_CHECK(hipMemcpyAsync(d_job, &j, sizeof(remesh::Job), hipMemcpyHostToDevice,
                                ctx->stream));
    remesh::Launch l;
    l.njobs = 1;
    l.max_panels = npanels;
    l.max_points = grid ? n : j.nx;
    l.max_rows = np;
    l.max_nx = j.nx;
    FS_TRY(remesh::clear_status(d_job, l, ctx->stream));
    FS_TRY(remesh::prepare_old_meshes(d_job, l, ctx->stream));
    if (grid) {
        FS_TRY(remesh::interp_grids(d_job, l, ctx->stream));
    } else {
        FS_TRY(remesh::interp_points(d_job, l, ctx->stream));
    }
    FS_HIP_CHECK(hipMemcpyAsync(f_host, j.f, sizeof(double) * n, hipMemcpyDeviceToHost,
                                ctx->stream));
    return read_status(ctx, j.status);
}

}  // namespace
}  // namespace fs

using namespace fs;

extern "C" {

int fs_remesh_interp_grid(fs_ctx* ctx, const int32_t* panels, int64_t n_panels,
                          const double* old_xs, const double* old_ps, const double* old_fs,
                          int64_t n_old, const double* new_xs, const double* new_ps, int nx,
                          int np, double x_min, double x_max, int periodic, double f_beyond,
                          double* f_out) {
    if (!ctx || !panels || !old_xs || !old_ps || !old_fs || !new_xs || !new_ps || !f_out)
        return fail(FS_E_INVALID, "fs_remesh_interp_grid: null argument");
    if (nx < 1 || np < 1 || (int64_t)nx * np > INT32_MAX || n_panels > INT32_MAX ||
        n_old > INT32_MAX)
        return fail(FS_E_INVALID, "fs_remesh_interp_grid: sizes out of range");
    if (nx > remesh::kMaxLine || np > remesh::kMaxLine)
        return fail(FS_E_INVALID, "fs_remesh_interp_grid: at most %d points per row and column",
                    remesh::kMaxLine);
    if (!(x_max > x_min)) return fail(FS_E_INVALID, "fs_remesh_interp_grid: x_max <= x_min");
    FS_TRY(check_panels(panels, n_panels, n_old));
    const int n = nx * np;
    // rows: points grouped by p (ascending), each row in input order
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return new_ps[a] < new_ps[b]; });
    for (int r = 0; r < np; ++r) {
        const double p0 = new_ps[order[(size_t)r * nx]];
        for (int c = 1; c < nx; ++c)
            if (new_ps[order[(size_t)r * nx + c]] != p0)
                return fail(FS_E_INVALID, "fs_remesh_interp_grid: points are not %d rows of %d",
                            np, nx);
    }
    FS_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t bn = align_up(sizeof(double) * n, 256);
    const size_t brow = align_up(sizeof(int32_t) * n, 256);
    Staged st;
    char* extra = nullptr;
    FS_TRY(stage_old_mesh(ctx, panels, n_panels, old_xs, old_ps, old_fs, n_old,
                          3 * bn + brow, st, extra));
    remesh::Job j{};
    j.rec = st.rec;
    j.npanels = (int)n_panels;
    j.ox = st.ox;
    j.op = st.op;
    j.of = st.of;
    double* d_x = reinterpret_cast<double*>(extra);
    double* d_p = reinterpret_cast<double*>(extra + bn);
    j.f = reinterpret_cast<double*>(extra + 2 * bn);
    int32_t* d_rows = reinterpret_cast<int32_t*>(extra + 3 * bn);
    FS_HIP_CHECK(hipMemcpyAsync(d_x, new_xs, sizeof(double) * n, hipMemcpyHostToDevice,
                                ctx->stream));
    FS_HIP_CHECK(hipMemcpyAsync(d_p, new_ps, sizeof(double) * n, hipMemcpyHostToDevice,
                                ctx->stream));
    FS_HIP_CHECK(hipMemcpyAsync(d_rows, order.data(), sizeof(int32_t) * n,
                                hipMemcpyHostToDevice, ctx->stream));
    j.x = d_x;
    j.p = d_p;
    j.rows = d_rows;
    j.nx = nx;
    j.np = np;
    j.Lx = x_max - x_min;
    j.periodic = periodic != 0;
    j.f_beyond = f_beyond;
    return run_job(ctx, j, (int)n_panels, n, np, true, f_out);
}

int fs_remesh_interp_points(fs_ctx* ctx, const int32_t* panels, int64_t n_panels,
                            const double* old_xs, const double* old_ps, const double* old_fs,
                            int64_t n_old, const double* x, const double* p, int64_t n,
                            double x_min, double x_max, int periodic, double f_beyond,
                            double* f_out) {
    if (!ctx || !panels || !old_xs || !old_ps || !old_fs || (n > 0 && (!x || !p || !f_out)))
        return fail(FS_E_INVALID, "fs_remesh_interp_points: null argument");
    if (n < 0 || n > INT32_MAX || n_panels > INT32_MAX || n_old > INT32_MAX)
        return fail(FS_E_INVALID, "fs_remesh_interp_points: sizes out of range");
    if (!(x_max > x_min)) return fail(FS_E_INVALID, "fs_remesh_interp_points: x_max <= x_min");
    FS_TRY(check_panels(panels, n_panels, n_old));
    if (n == 0) return FS_OK;
    FS_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t bn = align_up(sizeof(double) * n, 256);
    Staged st;
    char* extra = nullptr;
    FS_TRY(stage_old_mesh(ctx, panels, n_panels, old_xs, old_ps, old_fs, n_old, 3 * bn, st,
                          extra));
    remesh::Job j{};
    j.rec = st.rec;
    j.npanels = (int)n_panels;
    j.ox = st.ox;
    j.op = st.op;
    j.of = st.of;
    double* d_x = reinterpret_cast<double*>(extra);
    double* d_p = reinterpret_cast<double*>(extra + bn);
    j.f = reinterpret_cast<double*>(extra + 2 * bn);
    FS_HIP_CHECK(hipMemcpyAsync(d_x, x, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    FS_HIP_CHECK(hipMemcpyAsync(d_p, p, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    j.x = d_x;
    j.p = d_p;
    j.nx = (int)n;
    j.np = 1;
    j.Lx = x_max - x_min;
    j.periodic = periodic != 0;
    j.f_beyond = f_beyond;
    return run_job(ctx, j, (int)n_panels, (int)n, 1, false, f_out);
}

}  // extern "C"
