// The float64 assembly of a segment's result on the device: the distance-prior weights of every template
// over the points it reaches, the log-sum-exp over those points and the template average ln B.
//
//   exm_distance_weights   src/exmarg/search/pe.py:260-286 (distance_weights) for every template at once
//                          (pe.py:321-327): the weights w_j with sum_j w_j Lambda(y_j) = integral p(y)
//                          Lambda(y) dy for Lambda piecewise linear in x = ln y
//   exm_plan_set_weights   the weights of a plan's reads (ln w per read, cell weight per template)
//   exm_plan_gather_lse    exm_plan_gather followed by ln Lambda_m = LSE_k (ln w_km + lamN_km) (pe.py:333-336)
//                          and ln B = ln sum_m c_m e^{ln Lambda_m} over the finite entries (pe.py:135-142)
//
// Every sum runs in a fixed order: over a template's points in row order (the reference's node order, the
// boundary point last), over the templates of a segment in a fixed tree, so results do not depend on the
// launch.
#include <cmath>
#include <vector>

#include "exm_common.h"

#pragma clang fp contract(off)

namespace exm {

constexpr int DW_MAX_ROWS = 64;
constexpr int DW_POINTS = 401;    // trapezoid points per interval of a tabulated density

// the weights of one interval [a, b] in x = ln y on its two ends: left multiplies Lambda(a), right
// Lambda(b).  Volume prior p(y) = 3 ymin^3 y^-4 (density 3 ymin^3 e^{-3x} in x) integrated exactly;
// a tabulated density p(x) on lnx (ascending, zero outside) by the trapezoid rule on DW_POINTS points.
__device__ void interval_weights(double a, double b, int prior, double ymin, const double* __restrict__ lnx,
                                 const double* __restrict__ px, int ntab, int& j, double& left, double& right) {
    const double h = b - a;
    if (prior == 0) {
        const double c = 3.0 * ymin * ymin * ymin, k = 3.0;
        const double ea = exp(-k * a), eb = exp(-k * b);
        const double I0 = (ea - eb) / k;
        const double I1 = (a * ea - b * eb) / k + (ea - eb) / (k * k);
        left = c * (b * I0 - I1) / h;
        right = c * (I1 - a * I0) / h;
        return;
    }
    // x_i = a + i (b - a) / 400 (the last point exactly b); the density interpolated linearly with a
    // pointer j (lnx[j] <= x < lnx[j + 1]) that only moves up: the points of a template ascend
    const double step = (b - a) / (double)(DW_POINTS - 1);
    double sl = 0.0, sr = 0.0, xp = 0.0, gl = 0.0, gr = 0.0;
    for (int i = 0; i < DW_POINTS; ++i) {
        const double x = i == DW_POINTS - 1 ? b : a + (double)i * step;
        double f;
        if (x < lnx[0] || x > lnx[ntab - 1]) {
            f = 0.0;
        } else if (x == lnx[ntab - 1]) {
            f = px[ntab - 1];
        } else {
            while (j + 2 < ntab && lnx[j + 1] <= x) ++j;
            const double slope = (px[j + 1] - px[j]) / (lnx[j + 1] - lnx[j]);
            f = slope * (x - lnx[j]) + px[j];
        }
        const double yl = f * (b - x) / h, yr = f * (x - a) / h;
        if (i > 0) {
            const double dx = x - xp;
            sl += dx * (gl + yl) / 2.0;
            sr += dx * (gr + yr) / 2.0;
        }
        xp = x;
        gl = yl;
        gr = yr;
    }
    left = sl;
    right = sr;
}

// lnW (nrow, nm) of every template: the rows with y > 0 sorted by y (stable), the interval weights
// summed onto their ends, ln of the sums; -inf on the other rows
__global__ void k_distance_weights(const double* __restrict__ yN, int nrow, int nm, int prior, double ymin,
                                   const double* __restrict__ lnx, const double* __restrict__ px, int ntab,
                                   double* __restrict__ lnW) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nm) return;
    double y[DW_MAX_ROWS], w[DW_MAX_ROWS];
    int row[DW_MAX_ROWS];
    int cnt = 0;
    for (int r = 0; r < nrow; ++r) {
        const double v = yN[(int64_t)r * nm + m];
        lnW[(int64_t)r * nm + m] = -INFINITY;
        if (!(v > 0.0)) continue;
        int i = cnt++;
        while (i > 0 && y[i - 1] > v) {   // insertion keeps equal values in row order
            y[i] = y[i - 1];
            row[i] = row[i - 1];
            --i;
        }
        y[i] = v;
        row[i] = r;
    }
    for (int i = 0; i < cnt; ++i) w[i] = 0.0;
    int j = 0;
    for (int i = 0; i + 1 < cnt; ++i) {
        double left, right;
        interval_weights(log(y[i]), log(y[i + 1]), prior, ymin, lnx, px, ntab, j, left, right);
        w[i] += left;
        w[i + 1] += right;
    }
    for (int i = 0; i < cnt; ++i) lnW[(int64_t)row[i] * nm + m] = log(w[i]);
}

// L[m] = LSE over the reads of template m (positions pos[off[m] .. off[m+1]) of the table-sorted values,
// in the caller's read order) of lnw + val
__global__ void k_node_lse(const double* __restrict__ val, const double* __restrict__ lnw,
                           const int32_t* __restrict__ off, const int32_t* __restrict__ pos, int nm,
                           double* __restrict__ L) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nm) return;
    const int b = off[m], e = off[m + 1];
    double mx = -INFINITY;
    for (int i = b; i < e; ++i) {
        const int p = pos[i];
        mx = fmax(mx, lnw[p] + val[p]);
    }
    if (mx == -INFINITY) {
        L[m] = -INFINITY;
        return;
    }
    double s = 0.0;
    for (int i = b; i < e; ++i) {
        const int p = pos[i];
        s += exp(lnw[p] + val[p] - mx);
    }
    L[m] = mx + log(s);
}

constexpr int LM_THREADS = 256;

// ln_b[g] = ln sum c_m e^{L_m} over the templates seg[g] .. seg[g+1] with finite L and c > 0 (-inf if
// none); one workgroup per segment, a fixed strided split and tree
__global__ void __launch_bounds__(LM_THREADS) k_bank_mean(const double* __restrict__ L, const double* __restrict__ c,
                                                         const int32_t* __restrict__ seg, double* __restrict__ ln_b) {
    __shared__ double red[LM_THREADS];
    const int g = blockIdx.x, t = threadIdx.x;
    const int b = seg[g], e = seg[g + 1];
    double mx = -INFINITY;
    for (int m = b + t; m < e; m += LM_THREADS)
        if (isfinite(L[m]) && c[m] > 0.0) mx = fmax(mx, L[m]);
    red[t] = mx;
    __syncthreads();
    for (int k = LM_THREADS / 2; k > 0; k >>= 1) {
        if (t < k) red[t] = fmax(red[t], red[t + k]);
        __syncthreads();
    }
    mx = red[0];
    __syncthreads();
    double s = 0.0;
    if (mx != -INFINITY)
        for (int m = b + t; m < e; m += LM_THREADS)
            if (isfinite(L[m]) && c[m] > 0.0) s += c[m] * exp(L[m] - mx);
    red[t] = s;
    __syncthreads();
    for (int k = LM_THREADS / 2; k > 0; k >>= 1) {
        if (t < k) red[t] += red[t + k];
        __syncthreads();
    }
    if (t == 0) ln_b[g] = mx == -INFINITY ? -INFINITY : mx + log(red[0]);
}

void lse_layout(int64_t njobs, int n_tmpl, const int32_t* m, const std::vector<int64_t>& order, const double* lnw,
                LseLayout* out) {
    std::vector<int64_t> inv(njobs);
    for (int64_t i = 0; i < njobs; ++i) inv[order[i]] = i;
    out->off.assign(n_tmpl + 1, 0);
    for (int64_t e = 0; e < njobs; ++e) ++out->off[m[e] + 1];
    for (int k = 0; k < n_tmpl; ++k) out->off[k + 1] += out->off[k];
    out->pos.resize(njobs);
    std::vector<int32_t> fill(out->off.begin(), out->off.end() - 1);
    for (int64_t e = 0; e < njobs; ++e) out->pos[fill[m[e]]++] = (int32_t)inv[e];
    out->lnw.resize(njobs);
    for (int64_t e = 0; e < njobs; ++e) out->lnw[inv[e]] = lnw[e];
}

size_t lse_device_bytes(int64_t njobs, int64_t n_tmpl, int nseg) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    return al(4 * (size_t)(n_tmpl + 1)) + al(4 * (size_t)njobs) + al(8 * (size_t)njobs) + al(8 * (size_t)n_tmpl) +
           al(4 * (size_t)(nseg + 1)) + al(8 * (size_t)n_tmpl) + al(8 * (size_t)nseg);
}

void LseDevice::carve(void* base, int64_t njobs, int64_t n_tmpl, int nseg_) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    char* p = (char*)base;
    off = (int32_t*)p;
    p += al(4 * (size_t)(n_tmpl + 1));
    pos = (int32_t*)p;
    p += al(4 * (size_t)njobs);
    lnw = (double*)p;
    p += al(8 * (size_t)njobs);
    cell = (double*)p;
    p += al(8 * (size_t)n_tmpl);
    seg = (int32_t*)p;
    p += al(4 * (size_t)(nseg_ + 1));
    L = (double*)p;
    p += al(8 * (size_t)n_tmpl);
    ln_b = (double*)p;
    nm = n_tmpl;
    nseg = nseg_;
}

int LseDevice::upload(const LseLayout& lay, const double* cellw, const int32_t* segs, hipStream_t s) const {
    EXM_HIP(hipMemcpyAsync(off, lay.off.data(), 4 * lay.off.size(), hipMemcpyHostToDevice, s));
    EXM_HIP(hipMemcpyAsync(pos, lay.pos.data(), 4 * lay.pos.size(), hipMemcpyHostToDevice, s));
    EXM_HIP(hipMemcpyAsync(lnw, lay.lnw.data(), 8 * lay.lnw.size(), hipMemcpyHostToDevice, s));
    EXM_HIP(hipMemcpyAsync(cell, cellw, 8 * (size_t)nm, hipMemcpyHostToDevice, s));
    EXM_HIP(hipMemcpyAsync(seg, segs, 4 * (size_t)(nseg + 1), hipMemcpyHostToDevice, s));
    return EXM_OK;
}

int LseDevice::launch(const double* val, hipStream_t s) const {
    if (nm > 0) {
        hipLaunchKernelGGL(k_node_lse, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, s, val, lnw, off, pos,
                           (int)nm, L);
        EXM_HIP(hipGetLastError());
    }
    if (nseg > 0) {
        hipLaunchKernelGGL(k_bank_mean, dim3((unsigned)nseg), dim3(LM_THREADS), 0, s, L, cell, seg, ln_b);
        EXM_HIP(hipGetLastError());
    }
    return EXM_OK;
}

int check_node_weights(int n_tmpl, int64_t njobs, const double* lnw, const double* cellw) {
    for (int64_t e = 0; e < njobs; ++e)
        EXM_CHECK(!std::isnan(lnw[e]) && lnw[e] != INFINITY, EXM_ERR_ARG, "ln w of read %lld is %g", (long long)e,
                  lnw[e]);
    for (int k = 0; k < n_tmpl; ++k)
        EXM_CHECK(std::isfinite(cellw[k]) && cellw[k] >= 0.0, EXM_ERR_ARG, "cell weight %d is %g", k, cellw[k]);
    return EXM_OK;
}

}  // namespace exm

using namespace exm;

extern "C" int exm_distance_weights(const double* yN, int nrow, int nm, int prior, double ymin, const double* lnx,
                                    const double* px, int ntab, double* lnW) {
    std::lock_guard<std::recursive_mutex> lock(ctx().mu);
    EXM_TRY(ensure_init());
    EXM_CHECK(yN && lnW, EXM_ERR_ARG, "NULL argument");
    EXM_CHECK(nrow >= 2 && nrow <= DW_MAX_ROWS && nm > 0, EXM_ERR_ARG, "need 2 <= nrow <= %d and nm > 0 (got %d, %d)",
              DW_MAX_ROWS, nrow, nm);
    EXM_CHECK(prior == 0 || prior == 1, EXM_ERR_ARG, "prior must be 0 (volume) or 1 (tabulated), got %d", prior);
    EXM_CHECK(prior == 0 || (lnx && px && ntab >= 2), EXM_ERR_ARG, "a tabulated prior needs its density table");
    EXM_CHECK(ymin > 0.0 && std::isfinite(ymin), EXM_ERR_ARG, "ymin must be positive (got %g)", ymin);
    for (int m = 0; m < nm; ++m) {
        int cnt = 0;
        for (int r = 0; r < nrow; ++r) cnt += yN[(int64_t)r * nm + m] > 0.0;
        EXM_CHECK(cnt >= 2, EXM_ERR_ARG, "template %d reaches fewer than two points", m);
    }
    if (prior == 1)
        for (int i = 0; i + 1 < ntab; ++i)
            EXM_CHECK(lnx[i] < lnx[i + 1], EXM_ERR_ARG, "the density's abscissae must ascend (index %d)", i);
    const hipStream_t s = ctx().stream;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t ybytes = al(8 * (size_t)nrow * nm), tbytes = prior == 1 ? al(8 * (size_t)ntab) : 0;
    const size_t total = 2 * ybytes + 2 * tbytes;
    void* buf = nullptr;
    EXM_TRY(dev_alloc(&buf, total));
    double* dy = (double*)buf;
    double* dw = (double*)((char*)buf + ybytes);
    double* dx = prior == 1 ? (double*)((char*)buf + 2 * ybytes) : nullptr;
    double* dp = prior == 1 ? (double*)((char*)buf + 2 * ybytes + tbytes) : nullptr;
    int rc = EXM_OK;
    if (hipMemcpyAsync(dy, yN, 8 * (size_t)nrow * nm, hipMemcpyHostToDevice, s) != hipSuccess ||
        (prior == 1 && (hipMemcpyAsync(dx, lnx, 8 * (size_t)ntab, hipMemcpyHostToDevice, s) != hipSuccess ||
                        hipMemcpyAsync(dp, px, 8 * (size_t)ntab, hipMemcpyHostToDevice, s) != hipSuccess))) {
        set_error("distance weights: upload failed");
        rc = EXM_ERR_HIP;
    }
    if (rc == EXM_OK) {
        hipLaunchKernelGGL(k_distance_weights, dim3((unsigned)((nm + 63) / 64)), dim3(64), 0, s, dy, nrow, nm, prior,
                           ymin, dx, dp, prior == 1 ? ntab : 0, dw);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(lnW, dw, 8 * (size_t)nrow * nm, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            set_error("distance weights: kernel or copy failed");
            rc = EXM_ERR_HIP;
        }
    }
    dev_free(buf, total);
    return rc;
}

extern "C" int exm_plan_set_weights(exm_handle hplan, const double* lnw, const double* cellw) {
    std::lock_guard<std::recursive_mutex> lock(ctx().mu);
    EXM_TRY(ensure_init());
    Plan* plan = get<Plan>(hplan, K_PLAN);
    EXM_CHECK(plan != nullptr, EXM_ERR_HANDLE, "handle %llu is not a Plan", (unsigned long long)hplan);
    EXM_CHECK(lnw && cellw, EXM_ERR_ARG, "NULL argument");
    const int64_t njobs = plan->js.njobs;
    EXM_TRY(check_node_weights(plan->n_tmpl, njobs, lnw, cellw));
    plan->h_lnw.assign(lnw, lnw + njobs);
    plan->h_cellw.assign(cellw, cellw + plan->n_tmpl);
    LseLayout lay;
    lse_layout(njobs, plan->n_tmpl, plan->h_m.data(), plan->js.order, lnw, &lay);
    const size_t need = lse_device_bytes(njobs, plan->n_tmpl, 1);
    if (need > plan->w_bytes) {
        if (plan->d_w) dev_free(plan->d_w, plan->w_bytes);
        plan->d_w = nullptr;
        plan->w_bytes = 0;
        EXM_TRY(dev_alloc(&plan->d_w, need));
        plan->w_bytes = need;
    }
    plan->lse.carve(plan->d_w, njobs, plan->n_tmpl, 1);
    const int32_t segs[2] = {0, plan->n_tmpl};
    const hipStream_t s = ctx().stream;
    EXM_TRY(plan->lse.upload(lay, cellw, segs, s));
    EXM_HIP(hipStreamSynchronize(s));
    plan->has_weights = true;
    return EXM_OK;
}

extern "C" int exm_plan_gather_lse(exm_handle hplan, exm_handle htraj, exm_handle hstore, exm_handle htimes,
                                   double* ln_lambda, double* ln_b) {
    std::lock_guard<std::recursive_mutex> lock(ctx().mu);
    EXM_TRY(ensure_init());
    Plan* plan = get<Plan>(hplan, K_PLAN);
    Traj* traj = get<Traj>(htraj, K_TRAJ);
    Store* st = get<Store>(hstore, K_STORE);
    EXM_CHECK(plan && traj && st, EXM_ERR_HANDLE, "invalid plan, trajectory or store handle");
    EXM_CHECK(ln_lambda && ln_b, EXM_ERR_ARG, "NULL argument");
    EXM_CHECK(plan->has_weights, EXM_ERR_STATE, "plan has no node weights (exm_plan_set_weights)");
    Times* tm = nullptr;
    EXM_TRY(check_times(traj, htimes, &tm));
    EXM_CHECK(plan->n_tmpl == traj->n_tmpl, EXM_ERR_ARG, "plan is for %d templates, trajectories for %d", plan->n_tmpl,
              traj->n_tmpl);
    const int64_t njobs = plan->js.njobs;
    double* dout = nullptr;
    EXM_TRY(scratch(1, sizeof(double) * (size_t)njobs, (void**)&dout));
    EXM_TRY(launch_gather(traj, st, plan->js, tm, dout));
    const hipStream_t s = ctx().stream;
    EXM_TRY(plan->lse.launch(dout, s));
    EXM_HIP(hipMemcpyAsync(ln_lambda, plan->lse.L, sizeof(double) * (size_t)plan->n_tmpl, hipMemcpyDeviceToHost, s));
    EXM_HIP(hipMemcpyAsync(ln_b, plan->lse.ln_b, sizeof(double), hipMemcpyDeviceToHost, s));
    EXM_HIP(hipStreamSynchronize(s));
    return EXM_OK;
}
