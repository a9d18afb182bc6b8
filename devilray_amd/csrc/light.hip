// Emissive-triangle light sampler: a Walker alias table built on the device.
//
// Weights follow the reference's light weighting (src/Renderer.cpp:113-152:
// object weight = surface area x luminance(pi * emission); triangles of an
// object by area, src/GpuTris.cpp:7-27), flattened to one table over emissive
// triangles: w_t = area_t x luminance(E_t), both in FP64.
//
// Construction is the sweeping alias method (cf. Huebschle-Schneider &
// Sanders, "Parallel Weighted Random Sampling", 2022) in exact integer
// arithmetic, so every order of summation gives the same table:
//   q_i = floor(w_i / w_max * 2^S), Q = sum q_i, with S chosen so that every
//   prefix sum below fits in 63 bits;
//   item i is light if n q_i < Q (deficit Q - n q_i), heavy otherwise (excess
//   n q_i - Q); lights and heavies keep index order;
//   D_i = inclusive prefix of light deficits, E_j = inclusive prefix of heavy
//   excesses.
// A sequential sweep that lets the current heavy fill lights in order, and
// turns a heavy into its own (partial) bucket as soon as its remaining weight
// drops below one bucket, has remaining weight Q + E_j - D_i after serving
// light i with heavy j. Hence, in closed form and in parallel:
//   light i:  alias = min{ j : E_j >= D_{i-1} } (the last heavy at most),
//             prob  = n q_i / Q
//   heavy j (not the last): with i* = min{ i : D_i > E_j },
//             prob = (Q - (D_{i*} - E_j)) / Q, alias = heavy j+1;
//             no such i: prob = 1
//   last heavy: prob = 1.
// Selection probability of item i is q_i / Q; light_pdf_area[t] = that / area_t.
#include "dr_internal.hpp"

#include <cmath>

namespace dr {

namespace {

constexpr int kT = 256;

inline uint32_t blocks(uint32_t n) { return (n + kT - 1) / kT; }

__device__ __forceinline__ double luminance(const float *e)
{
    return 0.2126 * (double)e[0] + 0.7152 * (double)e[1] + 0.0722 * (double)e[2];
}

__global__ __launch_bounds__(kT) void k_emit_flags(const float *__restrict__ verts, const uint32_t *__restrict__ idx,
                                                   const uint32_t *__restrict__ tri_mat,
                                                   const dr_material *__restrict__ mats, uint32_t n,
                                                   uint32_t *__restrict__ flag, double *__restrict__ weight,
                                                   double *__restrict__ area)
{
    const uint32_t t = blockIdx.x * kT + threadIdx.x;
    if (t >= n) return;
    const dr_material m = mats[tri_mat[t]];
    const double lum = luminance(m.emission);
    double a3[3][3];
    for (int k = 0; k < 3; ++k) {
        const uint32_t vi = idx[3 * (size_t)t + k];
        for (int c = 0; c < 3; ++c) a3[k][c] = (double)verts[3 * (size_t)vi + c];
    }
    const double e1[3] = {a3[1][0] - a3[0][0], a3[1][1] - a3[0][1], a3[1][2] - a3[0][2]};
    const double e2[3] = {a3[2][0] - a3[0][0], a3[2][1] - a3[0][1], a3[2][2] - a3[0][2]};
    const double cx = e1[1] * e2[2] - e1[2] * e2[1];
    const double cy = e1[2] * e2[0] - e1[0] * e2[2];
    const double cz = e1[0] * e2[1] - e1[1] * e2[0];
    const double ar = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
    const bool em = m.kind == DR_MAT_LAMBERT && lum > 0.0 && ar > 0.0;
    flag[t] = em ? 1u : 0u;
    weight[t] = em ? ar * lum : 0.0;
    area[t] = ar;
}

__global__ __launch_bounds__(kT) void k_compact_emitters(const uint32_t *__restrict__ flag,
                                                         const uint32_t *__restrict__ pos, uint32_t n,
                                                         const double *__restrict__ weight,
                                                         uint32_t *__restrict__ emit_tri, double *__restrict__ ew,
                                                         unsigned long long *__restrict__ wmax_bits)
{
    const uint32_t t = blockIdx.x * kT + threadIdx.x;
    if (t >= n || !flag[t]) return;
    const uint32_t e = pos[t];
    emit_tri[e] = t;
    ew[e] = weight[t];
    // positive doubles order like their bit patterns
    atomicMax(wmax_bits, (unsigned long long)__double_as_longlong(weight[t]));
}

__global__ __launch_bounds__(kT) void k_quantise(const double *__restrict__ ew, uint32_t ne,
                                                 const unsigned long long *__restrict__ wmax_bits, double scale,
                                                 uint64_t *__restrict__ q, unsigned long long *__restrict__ qsum)
{
    const uint32_t e = blockIdx.x * kT + threadIdx.x;
    if (e >= ne) return;
    const double wmax = __longlong_as_double((long long)*wmax_bits);
    const uint64_t v = (uint64_t)floor(ew[e] / wmax * scale);
    q[e] = v;
    atomicAdd(qsum, (unsigned long long)v);
}

// lights / heavies: flags for compaction, values = deficit / excess
__global__ __launch_bounds__(kT) void k_classify(const uint64_t *__restrict__ q, uint32_t ne,
                                                 const unsigned long long *__restrict__ qsum,
                                                 uint32_t *__restrict__ is_light, uint32_t *__restrict__ is_heavy)
{
    const uint32_t e = blockIdx.x * kT + threadIdx.x;
    if (e >= ne) return;
    const uint64_t Q = *qsum;
    const bool light = (uint64_t)ne * q[e] < Q;
    is_light[e] = light ? 1u : 0u;
    is_heavy[e] = light ? 0u : 1u;
}

__global__ __launch_bounds__(kT) void k_split(const uint64_t *__restrict__ q, uint32_t ne,
                                              const unsigned long long *__restrict__ qsum,
                                              const uint32_t *__restrict__ is_light, const uint32_t *__restrict__ lpos,
                                              const uint32_t *__restrict__ hpos, uint32_t *__restrict__ light_item,
                                              uint32_t *__restrict__ heavy_item, uint64_t *__restrict__ D,
                                              uint64_t *__restrict__ E)
{
    const uint32_t e = blockIdx.x * kT + threadIdx.x;
    if (e >= ne) return;
    const uint64_t Q = *qsum;
    const uint64_t nq = (uint64_t)ne * q[e];
    if (is_light[e]) {
        light_item[lpos[e]] = e;
        D[lpos[e]] = Q - nq;
    } else {
        heavy_item[hpos[e]] = e;
        E[hpos[e]] = nq - Q;
    }
}

// exclusive -> inclusive prefix: value + own term
__global__ __launch_bounds__(kT) void k_make_inclusive(uint64_t *__restrict__ pre, const uint64_t *__restrict__ own,
                                                       uint32_t n)
{
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i < n) pre[i] += own[i];
}

struct AliasEntry {
    float p_a, pdf_a, pdf_b;
    int32_t a, b;
};
static_assert(sizeof(AliasEntry) == 20, "alias entry is 20 B");

__device__ __forceinline__ uint32_t lower_bound_ge(const uint64_t *a, uint32_t n, uint64_t x)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] >= x) hi = mid; else lo = mid + 1;
    }
    return lo;
}

__device__ __forceinline__ uint32_t upper_bound_gt(const uint64_t *a, uint32_t n, uint64_t x)
{
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] > x) hi = mid; else lo = mid + 1;
    }
    return lo;
}

__global__ __launch_bounds__(kT) void k_alias_lights(const uint32_t *__restrict__ light_item, uint32_t na,
                                                     const uint64_t *__restrict__ D, const uint64_t *__restrict__ E,
                                                     const uint32_t *__restrict__ heavy_item, uint32_t nb,
                                                     const uint64_t *__restrict__ q, uint32_t ne,
                                                     const unsigned long long *__restrict__ qsum,
                                                     const uint32_t *__restrict__ emit_tri,
                                                     AliasEntry *__restrict__ table)
{
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i >= na) return;
    const uint64_t Q = *qsum;
    const uint64_t dprev = i ? D[i - 1] : 0ull;
    uint32_t j = lower_bound_ge(E, nb, dprev);
    if (j > nb - 1) j = nb - 1;
    const uint32_t e = light_item[i], h = heavy_item[j];
    AliasEntry en;
    en.p_a = (float)((double)((uint64_t)ne * q[e]) / (double)Q);
    en.pdf_a = (float)((double)q[e] / (double)Q);
    en.pdf_b = (float)((double)q[h] / (double)Q);
    en.a = emit_tri ? (int32_t)emit_tri[e] : (int32_t)e;
    en.b = emit_tri ? (int32_t)emit_tri[h] : (int32_t)h;
    table[e] = en;
}

__global__ __launch_bounds__(kT) void k_alias_heavies(const uint32_t *__restrict__ heavy_item, uint32_t nb,
                                                      const uint64_t *__restrict__ E, const uint64_t *__restrict__ D,
                                                      uint32_t na, const uint64_t *__restrict__ q,
                                                      const unsigned long long *__restrict__ qsum,
                                                      const uint32_t *__restrict__ emit_tri,
                                                      AliasEntry *__restrict__ table)
{
    const uint32_t j = blockIdx.x * kT + threadIdx.x;
    if (j >= nb) return;
    const uint64_t Q = *qsum;
    const uint32_t e = heavy_item[j];
    AliasEntry en;
    en.p_a = 1.0f;
    en.pdf_a = (float)((double)q[e] / (double)Q);
    en.pdf_b = 0.0f;
    en.a = emit_tri ? (int32_t)emit_tri[e] : (int32_t)e;
    en.b = -1;
    if (j + 1 < nb) {
        const uint32_t i = upper_bound_gt(D, na, E[j]);
        if (i < na) {
            const uint32_t h = heavy_item[j + 1];
            en.p_a = (float)((double)(Q - (D[i] - E[j])) / (double)Q);
            en.pdf_b = (float)((double)q[h] / (double)Q);
            en.b = emit_tri ? (int32_t)emit_tri[h] : (int32_t)h;
        }
    }
    table[e] = en;
}

__global__ __launch_bounds__(kT) void k_pdf_area(const uint32_t *__restrict__ emit_tri, uint32_t ne,
                                                 const uint64_t *__restrict__ q,
                                                 const unsigned long long *__restrict__ qsum,
                                                 const double *__restrict__ area, float *__restrict__ pdf_area)
{
    const uint32_t e = blockIdx.x * kT + threadIdx.x;
    if (e >= ne) return;
    const uint32_t t = emit_tri[e];
    pdf_area[t] = (float)((double)q[e] / (double)*qsum / area[t]);
}

__global__ __launch_bounds__(kT) void k_fill_f32(float *p, uint32_t n, float v)
{
    const uint32_t i = blockIdx.x * kT + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_zero_u64(unsigned long long *p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

__global__ __launch_bounds__(kT) void k_pdf_emit(uint32_t ne, const uint64_t *__restrict__ q,
                                                 const unsigned long long *__restrict__ qsum,
                                                 const double *__restrict__ area_e, float *__restrict__ pdf)
{
    const uint32_t e = blockIdx.x * kT + threadIdx.x;
    if (e >= ne) return;
    pdf[e] = (float)((double)q[e] / (double)*qsum / area_e[e]);
}

// ---- instance forests ------------------------------------------------------

// candidates: triangles of the concatenated scenes whose material emits
__global__ __launch_bounds__(kT) void k_emissive_material(const uint32_t *__restrict__ tri_mat,
                                                          const dr_material *__restrict__ mats, uint32_t n,
                                                          uint32_t *__restrict__ flag)
{
    const uint32_t t = blockIdx.x * kT + threadIdx.x;
    if (t >= n) return;
    const dr_material m = mats[tri_mat[t]];
    flag[t] = (m.kind == DR_MAT_LAMBERT && luminance(m.emission) > 0.0) ? 1u : 0u;
}

__global__ __launch_bounds__(kT) void k_compact_u32(const uint32_t *__restrict__ flag,
                                                    const uint32_t *__restrict__ pos, uint32_t n,
                                                    uint32_t *__restrict__ out)
{
    const uint32_t t = blockIdx.x * kT + threadIdx.x;
    if (t < n && flag[t]) out[pos[t]] = t;
}

// Per instance: its first candidate in the enumeration over all instances,
// and where its scene's candidates start in the compacted candidate list.
struct InstCand {
    uint64_t first;
    uint32_t cstart, pad;
};

// Candidate c of the enumeration (instances in order, each scene's
// candidates in triangle order, so instanced triangle ids ascend): its world
// triangle, area and weight exactly as k_emit_flags computes them for the
// flattened copy of the instances.
__global__ __launch_bounds__(kT) void k_instance_emitters(const InstCand *__restrict__ ic, uint32_t n_inst,
                                                          uint64_t n_cand, const uint32_t *__restrict__ cand,
                                                          const InstShade *__restrict__ inst,
                                                          const float *__restrict__ verts,
                                                          const uint32_t *__restrict__ idx,
                                                          const uint32_t *__restrict__ tri_mat,
                                                          const dr_material *__restrict__ mats,
                                                          uint32_t *__restrict__ flag, uint32_t *__restrict__ id,
                                                          double *__restrict__ weight, double *__restrict__ area)
{
    const uint64_t c = (uint64_t)blockIdx.x * kT + threadIdx.x;
    if (c >= n_cand) return;
    uint32_t lo = 0, hi = n_inst;  // the last instance whose first candidate is <= c
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ic[mid].first <= c) lo = mid; else hi = mid;
    }
    const InstShade r = inst[lo];
    const uint32_t ct = cand[ic[lo].cstart + (uint32_t)(c - ic[lo].first)];
    const dr_material m = mats[tri_mat[ct]];
    const double lum = luminance(m.emission);
    const float s[3] = {r.sx, r.sy, r.sz}, tr[3] = {r.tx, r.ty, r.tz};
    double a3[3][3];
    for (int k = 0; k < 3; ++k) {
        const uint32_t vi = idx[3 * (size_t)ct + k];
        for (int a = 0; a < 3; ++a) {
            const float w = verts[3 * (size_t)vi + a] * s[a] + tr[a];
            a3[k][a] = (double)w;
        }
    }
    const double e1[3] = {a3[1][0] - a3[0][0], a3[1][1] - a3[0][1], a3[1][2] - a3[0][2]};
    const double e2[3] = {a3[2][0] - a3[0][0], a3[2][1] - a3[0][1], a3[2][2] - a3[0][2]};
    const double cx = e1[1] * e2[2] - e1[2] * e2[1];
    const double cy = e1[2] * e2[0] - e1[0] * e2[2];
    const double cz = e1[0] * e2[1] - e1[1] * e2[0];
    const double ar = 0.5 * sqrt(cx * cx + cy * cy + cz * cz);
    const bool em = lum > 0.0 && ar > 0.0;
    flag[c] = em ? 1u : 0u;
    id[c] = r.id_delta + (ct - r.tri_off);
    weight[c] = em ? ar * lum : 0.0;
    area[c] = ar;
}

__global__ __launch_bounds__(kT) void k_compact_instance_emitters(const uint32_t *__restrict__ flag,
                                                                  const uint32_t *__restrict__ pos, uint32_t n,
                                                                  const uint32_t *__restrict__ id,
                                                                  const double *__restrict__ weight,
                                                                  const double *__restrict__ area,
                                                                  uint32_t *__restrict__ emit_tri,
                                                                  double *__restrict__ ew, double *__restrict__ area_e,
                                                                  unsigned long long *__restrict__ wmax_bits)
{
    const uint32_t c = blockIdx.x * kT + threadIdx.x;
    if (c >= n || !flag[c]) return;
    const uint32_t e = pos[c];
    emit_tri[e] = id[c];
    ew[e] = weight[c];
    area_e[e] = area[c];
    atomicMax(wmax_bits, (unsigned long long)__double_as_longlong(weight[c]));
}

// The alias table over ne emitters with weights ew (the largest one's bits in
// sc[0]); writes Q to sc[1] and the quantised weights to q. Entries name
// ids[e], or the emitter index e when ids is NULL.
dr_status alias_table(Scene *s, uint32_t ne, DevBuf &ew, DevBuf &q, const uint32_t *ids, hipStream_t stream)
{
    DevBuf is_light, is_heavy, lpos, hpos, light_item, heavy_item, D, E, own;
    dr_status st = DR_OK;
    auto need = [&](DevBuf &b, size_t bytes) { if (st == DR_OK) st = b.ensure(bytes); };
    need(q, 8ull * ne);
    need(is_light, 4ull * ne);
    need(is_heavy, 4ull * ne);
    need(lpos, 4ull * ne);
    need(hpos, 4ull * ne);
    need(light_item, 4ull * ne);
    need(heavy_item, 4ull * ne);
    need(D, 8ull * ne);
    need(E, 8ull * ne);
    need(own, 8ull * ne);
    need(s->alias, sizeof(AliasEntry) * (size_t)ne);
    if (st) return st;
    unsigned long long *sc = s->light_scalars.as<unsigned long long>();
    // scale 2^S with n^2 2^S < 2^62 so that all deficit / excess prefixes fit
    int L = 1;
    while ((1ull << L) < ne) ++L;
    const int S = std::min(52, 62 - 2 * L);
    k_quantise<<<blocks(ne), kT, 0, stream>>>(ew.as<double>(), ne, sc, std::ldexp(1.0, S), q.as<uint64_t>(), sc + 1);
    k_classify<<<blocks(ne), kT, 0, stream>>>(q.as<uint64_t>(), ne, sc + 1, is_light.as<uint32_t>(), is_heavy.as<uint32_t>());
    DR_HIP_TRY(hipGetLastError());
    DR_HIP_TRY(hipMemcpyAsync(lpos.ptr, is_light.ptr, 4ull * ne, hipMemcpyDeviceToDevice, stream));
    DR_HIP_TRY(hipMemcpyAsync(hpos.ptr, is_heavy.ptr, 4ull * ne, hipMemcpyDeviceToDevice, stream));
    if ((st = exclusive_scan_u32(s->ctx, lpos.as<uint32_t>(), ne, stream))) return st;
    if ((st = exclusive_scan_u32(s->ctx, hpos.as<uint32_t>(), ne, stream))) return st;
    uint32_t lp = 0, ll = 0;
    DR_HIP_TRY(hipMemcpyAsync(&lp, lpos.as<uint32_t>() + ne - 1, 4, hipMemcpyDeviceToHost, stream));
    DR_HIP_TRY(hipMemcpyAsync(&ll, is_light.as<uint32_t>() + ne - 1, 4, hipMemcpyDeviceToHost, stream));
    DR_HIP_TRY(hipStreamSynchronize(stream));
    const uint32_t na = lp + ll, nb = ne - na;
    k_split<<<blocks(ne), kT, 0, stream>>>(q.as<uint64_t>(), ne, sc + 1, is_light.as<uint32_t>(), lpos.as<uint32_t>(),
                                           hpos.as<uint32_t>(), light_item.as<uint32_t>(), heavy_item.as<uint32_t>(),
                                           D.as<uint64_t>(), E.as<uint64_t>());
    DR_HIP_TRY(hipGetLastError());
    if (na) {
        DR_HIP_TRY(hipMemcpyAsync(own.ptr, D.ptr, 8ull * na, hipMemcpyDeviceToDevice, stream));
        if ((st = exclusive_scan_u64(s->ctx, D.as<uint64_t>(), na, stream))) return st;
        k_make_inclusive<<<blocks(na), kT, 0, stream>>>(D.as<uint64_t>(), own.as<uint64_t>(), na);
    }
    DR_HIP_TRY(hipMemcpyAsync(own.ptr, E.ptr, 8ull * nb, hipMemcpyDeviceToDevice, stream));
    if ((st = exclusive_scan_u64(s->ctx, E.as<uint64_t>(), nb, stream))) return st;
    k_make_inclusive<<<blocks(nb), kT, 0, stream>>>(E.as<uint64_t>(), own.as<uint64_t>(), nb);
    AliasEntry *table = s->alias.as<AliasEntry>();
    if (na)
        k_alias_lights<<<blocks(na), kT, 0, stream>>>(light_item.as<uint32_t>(), na, D.as<uint64_t>(), E.as<uint64_t>(),
                                                      heavy_item.as<uint32_t>(), nb, q.as<uint64_t>(), ne, sc + 1, ids,
                                                      table);
    k_alias_heavies<<<blocks(nb), kT, 0, stream>>>(heavy_item.as<uint32_t>(), nb, E.as<uint64_t>(), D.as<uint64_t>(), na,
                                                   q.as<uint64_t>(), sc + 1, ids, table);
    DR_HIP_TRY(hipGetLastError());
    // the temporaries are freed on return: wait for the kernels that read them
    DR_HIP_TRY(hipStreamSynchronize(stream));
    return DR_OK;
}

}  // namespace

dr_status build_lights(Scene *s, hipStream_t stream)
{
    const uint32_t n = s->n_tris;
    if (n == 0) return fail(DR_E_INVALID, "scene has no triangles");
    DevBuf flag, pos, weight, ew, q;
    dr_status st = DR_OK;
    auto need = [&](DevBuf &b, size_t bytes) { if (st == DR_OK) st = b.ensure(bytes); };
    need(flag, 4ull * n);
    need(pos, 4ull * n);
    need(weight, 8ull * n);
    need(s->tri_area, 8ull * n);
    need(s->light_pdf_area, 4ull * n);
    need(s->light_scalars, 64);
    if (st) return st;
    unsigned long long *sc = s->light_scalars.as<unsigned long long>();  // [0] wmax bits, [1] Q
    k_zero_u64<<<1, 64, 0, stream>>>(sc, 8);
    k_emit_flags<<<blocks(n), kT, 0, stream>>>(s->verts.as<float>(), s->indices.as<uint32_t>(), s->tri_mat.as<uint32_t>(),
                                               s->materials.as<dr_material>(), n, flag.as<uint32_t>(),
                                               weight.as<double>(), s->tri_area.as<double>());
    k_fill_f32<<<blocks(n), kT, 0, stream>>>(s->light_pdf_area.as<float>(), n, 0.0f);
    DR_HIP_TRY(hipGetLastError());
    DR_HIP_TRY(hipMemcpyAsync(pos.ptr, flag.ptr, 4ull * n, hipMemcpyDeviceToDevice, stream));
    if ((st = exclusive_scan_u32(s->ctx, pos.as<uint32_t>(), n, stream))) return st;
    uint32_t last_pos = 0, last_flag = 0;
    DR_HIP_TRY(hipMemcpyAsync(&last_pos, pos.as<uint32_t>() + n - 1, 4, hipMemcpyDeviceToHost, stream));
    DR_HIP_TRY(hipMemcpyAsync(&last_flag, flag.as<uint32_t>() + n - 1, 4, hipMemcpyDeviceToHost, stream));
    DR_HIP_TRY(hipStreamSynchronize(stream));
    const uint32_t ne = last_pos + last_flag;
    s->n_emit = ne;
    s->lights_built = true;
    if (ne == 0) return DR_OK;

    need(s->emit_tri, 4ull * ne);
    need(ew, 8ull * ne);
    if (st) return st;
    k_compact_emitters<<<blocks(n), kT, 0, stream>>>(flag.as<uint32_t>(), pos.as<uint32_t>(), n, weight.as<double>(),
                                                     s->emit_tri.as<uint32_t>(), ew.as<double>(), sc);
    DR_HIP_TRY(hipGetLastError());
    if ((st = alias_table(s, ne, ew, q, s->emit_tri.as<uint32_t>(), stream))) return st;
    k_pdf_area<<<blocks(ne), kT, 0, stream>>>(s->emit_tri.as<uint32_t>(), ne, q.as<uint64_t>(), sc + 1,
                                              s->tri_area.as<double>(), s->light_pdf_area.as<float>());
    DR_HIP_TRY(hipGetLastError());
    unsigned long long qs = 0;
    DR_HIP_TRY(hipMemcpyAsync(&qs, sc + 1, 8, hipMemcpyDeviceToHost, stream));
    DR_HIP_TRY(hipStreamSynchronize(stream));
    s->light_q_total = qs;
    return DR_OK;
}

dr_status build_instance_lights(Scene *f, hipStream_t stream)
{
    const uint32_t nc = f->n_cat_tris;
    const uint32_t n_inst = f->n_inst;
    if (nc == 0 || n_inst == 0) return fail(DR_E_INVALID, "instance set has no triangles");
    if (f->cat_tri_off.empty() || f->inst_scene.size() != n_inst) return fail(DR_E_INVALID, "not an instance forest");
    const size_t n_sc = f->cat_tri_off.size() - 1;
    DevBuf flag, pos, cand, icand, cflag, cpos, cid, cweight, carea, ew, area_e, q;
    dr_status st = DR_OK;
    auto need = [&](DevBuf &b, size_t bytes) { if (st == DR_OK) st = b.ensure(bytes); };
    need(flag, 4ull * nc);
    need(pos, 4ull * nc);
    need(f->light_scalars, 64);
    if (st) return st;
    unsigned long long *sc = f->light_scalars.as<unsigned long long>();  // [0] wmax bits, [1] Q
    k_zero_u64<<<1, 64, 0, stream>>>(sc, 8);
    k_emissive_material<<<blocks(nc), kT, 0, stream>>>(f->tri_mat.as<uint32_t>(), f->materials.as<dr_material>(), nc,
                                                       flag.as<uint32_t>());
    DR_HIP_TRY(hipGetLastError());
    DR_HIP_TRY(hipMemcpyAsync(pos.ptr, flag.ptr, 4ull * nc, hipMemcpyDeviceToDevice, stream));
    if ((st = exclusive_scan_u32(f->ctx, pos.as<uint32_t>(), nc, stream))) return st;
    // scene k's candidates are [cstart[k], cstart[k + 1]) of the compacted list
    std::vector<uint32_t> cstart(n_sc + 1);
    uint32_t last_flag = 0;
    for (size_t k = 0; k < n_sc; ++k) {
        const uint32_t b = f->cat_tri_off[k];
        if (b < nc) DR_HIP_TRY(hipMemcpyAsync(&cstart[k], pos.as<uint32_t>() + b, 4, hipMemcpyDeviceToHost, stream));
    }
    DR_HIP_TRY(hipMemcpyAsync(&cstart[n_sc], pos.as<uint32_t>() + nc - 1, 4, hipMemcpyDeviceToHost, stream));
    DR_HIP_TRY(hipMemcpyAsync(&last_flag, flag.as<uint32_t>() + nc - 1, 4, hipMemcpyDeviceToHost, stream));
    DR_HIP_TRY(hipStreamSynchronize(stream));
    cstart[n_sc] += last_flag;
    for (size_t k = 0; k < n_sc; ++k)
        if (f->cat_tri_off[k] >= nc) cstart[k] = cstart[n_sc];  // scenes without triangles at the end
    f->n_emit = 0;
    f->lights_built = true;
    const uint32_t n_cand_cat = cstart[n_sc];
    if (n_cand_cat == 0) return DR_OK;
    // the enumeration over instances
    std::vector<InstCand> ic(n_inst);
    uint64_t total = 0;
    for (uint32_t i = 0; i < n_inst; ++i) {
        const uint32_t k = f->inst_scene[i];
        ic[i].first = total;
        ic[i].cstart = cstart[k];
        ic[i].pad = 0;
        total += cstart[k + 1] - cstart[k];
    }
    if (total == 0) return DR_OK;
    if (total > 0x7FFFFFFFull) return fail(DR_E_INVALID, "instances hold more than 2^31 - 1 emissive triangles");
    const uint32_t nt = (uint32_t)total;
    need(cand, 4ull * n_cand_cat);
    if (st) return st;
    k_compact_u32<<<blocks(nc), kT, 0, stream>>>(flag.as<uint32_t>(), pos.as<uint32_t>(), nc, cand.as<uint32_t>());
    DR_HIP_TRY(hipGetLastError());
    need(icand, sizeof(InstCand) * (size_t)n_inst);
    need(cflag, 4ull * nt);
    need(cpos, 4ull * nt);
    need(cid, 4ull * nt);
    need(cweight, 8ull * nt);
    need(carea, 8ull * nt);
    if (st) return st;
    DR_HIP_TRY(hipMemcpyAsync(icand.ptr, ic.data(), sizeof(InstCand) * (size_t)n_inst, hipMemcpyHostToDevice, stream));
    k_instance_emitters<<<blocks(nt), kT, 0, stream>>>(icand.as<InstCand>(), n_inst, nt, cand.as<uint32_t>(),
                                                       f->inst_shade.as<InstShade>(), f->verts.as<float>(),
                                                       f->indices.as<uint32_t>(), f->tri_mat.as<uint32_t>(),
                                                       f->materials.as<dr_material>(), cflag.as<uint32_t>(),
                                                       cid.as<uint32_t>(), cweight.as<double>(), carea.as<double>());
    DR_HIP_TRY(hipGetLastError());
    DR_HIP_TRY(hipMemcpyAsync(cpos.ptr, cflag.ptr, 4ull * nt, hipMemcpyDeviceToDevice, stream));
    if ((st = exclusive_scan_u32(f->ctx, cpos.as<uint32_t>(), nt, stream))) return st;
    uint32_t lp = 0, lf = 0;
    DR_HIP_TRY(hipMemcpyAsync(&lp, cpos.as<uint32_t>() + nt - 1, 4, hipMemcpyDeviceToHost, stream));
    DR_HIP_TRY(hipMemcpyAsync(&lf, cflag.as<uint32_t>() + nt - 1, 4, hipMemcpyDeviceToHost, stream));
    DR_HIP_TRY(hipStreamSynchronize(stream));
    const uint32_t ne = lp + lf;
    if (ne == 0) return DR_OK;
    need(f->emit_tri, 4ull * ne);
    need(f->emit_pdf, 4ull * ne);
    need(ew, 8ull * ne);
    need(area_e, 8ull * ne);
    if (st) return st;
    k_compact_instance_emitters<<<blocks(nt), kT, 0, stream>>>(cflag.as<uint32_t>(), cpos.as<uint32_t>(), nt,
                                                               cid.as<uint32_t>(), cweight.as<double>(),
                                                               carea.as<double>(), f->emit_tri.as<uint32_t>(),
                                                               ew.as<double>(), area_e.as<double>(), sc);
    DR_HIP_TRY(hipGetLastError());
    if ((st = alias_table(f, ne, ew, q, nullptr, stream))) return st;
    k_pdf_emit<<<blocks(ne), kT, 0, stream>>>(ne, q.as<uint64_t>(), sc + 1, area_e.as<double>(),
                                              f->emit_pdf.as<float>());
    DR_HIP_TRY(hipGetLastError());
    unsigned long long qs = 0;
    DR_HIP_TRY(hipMemcpyAsync(&qs, sc + 1, 8, hipMemcpyDeviceToHost, stream));
    DR_HIP_TRY(hipStreamSynchronize(stream));
    f->light_q_total = qs;
    f->n_emit = ne;
    return DR_OK;
}

}  // namespace dr
