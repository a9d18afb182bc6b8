/* Learned path guide: FP64 restatement of the network, its mixture head, the
 * training loss and the exact gradient (test infrastructure only).
 *
 * There is no reference counterpart: the reference samples Lambert bounces
 * from the cosine BSDF alone (include/tracing/PathGeneration.hpp:406-420).
 * The math pinned here:
 *   - vMF density kappa/(2 pi (1 - e^-2kappa)) e^(kappa (mu.w - 1)) (unit
 *     integral on the sphere: checked by quadrature in the tests);
 *   - W. Jakob, "Numerically stable sampling of the von Mises-Fisher
 *     distribution on S^2" (2012) for sampling (checked against the density
 *     and against the closed-form mean cosine coth(kappa) - 1/kappa);
 *   - loss sum_i w_i * [-log(q/2 + cos+/(2 pi)) - log(alpha q' + (1 - alpha)
 *     cos+/pi)]: the lobes learn the KL of a fixed half-and-half proposal,
 *     the selection probability that of the learned one-sample MIS pdf with
 *     the mixture q' held fixed; the gradient is derived by hand below and
 *     checked against central differences of or_guide_loss (which takes the
 *     frozen parameters of q' separately).
 * With emulate_bf16 the functions round where the device kernels hold BF16
 * (weights, input features, hidden activations, output and hidden deltas),
 * so the device can be compared at FP32-accumulation tolerance. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define NL 5
static const int G_IN[NL] = {32, 64, 64, 64, 64};
static const int G_OUT[NL] = {64, 64, 64, 64, 48};
#define KR_LO (-3.0)
#define KR_HI 7.0
#define R_MIN 1e-6
#define A_LO 0.1
#define A_SPAN 0.8
#define Q_MIX 0.5 /* mixture share in the lobes' training pdf */
#define ORPI 3.14159265358979323846

static int w_off(int l)
{
    int off = 0;
    for (int k = 0; k < l; ++k) off += G_OUT[k] * G_IN[k] + G_OUT[k];
    return off;
}
static int b_off(int l) { return w_off(l) + G_OUT[l] * G_IN[l]; }

static double bf16r(double x)
{
    float f = (float)x;
    uint32_t b;
    memcpy(&b, &f, 4);
    b = (b + 0x7FFFu + ((b >> 16) & 1u)) & 0xFFFF0000u;
    memcpy(&f, &b, 4);
    return (double)f;
}

/* output row of lobe j component c (0..2 mu, 3 log kappa, 4 logit) */
static int head_row(int lobe, int comp)
{
    const int g = lobe / 2;
    const int q = (lobe % 2) * 5 + comp;
    return (q / 4) * 16 + 4 * g + (q % 4);
}
#define SEL_ROW 34

/* per axis sin and cos of 2^(k+1) pi u, k = 0..3 (u in [0, 1] across the box),
 * from the same float reduction of u * 2^k to [0, 1) and the same sin/cos
 * polynomial as the device, so the features match bit for bit */
void or_guide_encode(const float bounds[6], const or_guide_sample *s, double feat[32])
{
    for (int d = 0; d < 3; ++d) {
        const float ext = bounds[3 + d] - bounds[d];
        float u = (s->pos[d] - bounds[d]) / ext;
        u = fminf(fmaxf(u, 0.0f), 1.0f);
        float scale = 1.0f;
        for (int k = 0; k < 4; ++k) {
            const float a = u * scale;
            const float t = a - floorf(a);
            float sn, cs;
            or_sincos_2pi(t, &sn, &cs);  /* the device's polynomial, bit for bit */
            feat[8 * d + 2 * k] = sn;
            feat[8 * d + 2 * k + 1] = cs;
            scale *= 2.0f;
        }
    }
    for (int k = 0; k < 3; ++k) {
        feat[24 + k] = s->normal[k];
        feat[27 + k] = s->wo[k];
    }
    feat[30] = 1.0;
    feat[31] = 0.0;
}

typedef struct {
    double mu[8][3], r[8], kr[8], kappa[8], pi[8], sig, alpha;
} head_t;

static void decode(const double *y, head_t *h)
{
    double wmax = -INFINITY;
    for (int j = 0; j < 8; ++j) {
        const double mx = y[head_row(j, 0)], my = y[head_row(j, 1)], mz = y[head_row(j, 2)];
        double r = sqrt(mx * mx + my * my + mz * mz);
        if (r < R_MIN) r = R_MIN;
        h->r[j] = r;
        h->mu[j][0] = mx / r;
        h->mu[j][1] = my / r;
        h->mu[j][2] = mz / r;
        h->kr[j] = y[head_row(j, 3)];
        const double kc = h->kr[j] < KR_LO ? KR_LO : (h->kr[j] > KR_HI ? KR_HI : h->kr[j]);
        h->kappa[j] = exp(kc);
        const double w = y[head_row(j, 4)];
        if (w > wmax) wmax = w;
    }
    double sum = 0.0;
    for (int j = 0; j < 8; ++j) {
        h->pi[j] = exp(y[head_row(j, 4)] - wmax);
        sum += h->pi[j];
    }
    for (int j = 0; j < 8; ++j) h->pi[j] /= sum;
    h->sig = 1.0 / (1.0 + exp(-y[SEL_ROW]));
    h->alpha = A_LO + A_SPAN * h->sig;
}

double or_vmf_pdf(const double mu[3], double kappa, const double w[3])
{
    const double c = mu[0] * w[0] + mu[1] * w[1] + mu[2] * w[2];
    return kappa / (2.0 * ORPI * -expm1(-2.0 * kappa)) * exp(kappa * (c - 1.0));
}

void or_vmf_sample(const double mu[3], double kappa, double x1, double x2, double out[3])
{
    double w = 1.0 + log(x1 + (1.0 - x1) * exp(-2.0 * kappa)) / kappa;
    if (w > 1.0) w = 1.0;
    if (w < -1.0) w = -1.0;
    const double st = sqrt(fmax(0.0, 1.0 - w * w));
    const double ph = 2.0 * ORPI * x2;
    /* orthonormal basis about mu (Duff et al. 2017) */
    const double sgn = mu[2] >= 0.0 ? 1.0 : -1.0;
    const double a = -1.0 / (sgn + mu[2]);
    const double b = mu[0] * mu[1] * a;
    const double t[3] = {1.0 + sgn * mu[0] * mu[0] * a, sgn * b, -sgn * mu[0]};
    const double s[3] = {b, sgn + mu[1] * mu[1] * a, -mu[1]};
    for (int k = 0; k < 3; ++k) out[k] = w * mu[k] + st * (cos(ph) * t[k] + sin(ph) * s[k]);
}

static double mixture(const head_t *h, const double w[3], double f[8])
{
    double q = 0.0;
    for (int j = 0; j < 8; ++j) {
        f[j] = or_vmf_pdf(h->mu[j], h->kappa[j], w);
        q += h->pi[j] * f[j];
    }
    return q;
}

/* activations of one sample: x[0] (32), x[1..4] (64), y (48) */
typedef struct {
    double x[NL][64];
    double y[48];
} acts_t;

static void forward_one(const double *P, const float bounds[6], const or_guide_sample *s, int emu, acts_t *a)
{
    double feat[32];
    or_guide_encode(bounds, s, feat);
    for (int k = 0; k < 32; ++k) a->x[0][k] = emu ? bf16r(feat[k]) : feat[k];
    for (int l = 0; l < NL; ++l) {
        const double *W = P + w_off(l);
        const double *b = P + b_off(l);
        const double *in = a->x[l];
        for (int o = 0; o < G_OUT[l]; ++o) {
            double z = 0.0;
            for (int i = 0; i < G_IN[l]; ++i) {
                const double wv = emu ? bf16r(W[o * G_IN[l] + i]) : W[o * G_IN[l] + i];
                z += wv * in[i];
            }
            z += b[o];
            if (l == NL - 1) {
                a->y[o] = z;
            } else {
                const double r = z > 0.0 ? z : 0.0;
                a->x[l + 1][o] = emu ? bf16r(r) : r;
            }
        }
    }
}

static void decode_out(const head_t *h, double *out)
{
    for (int j = 0; j < 8; ++j) {
        out[5 * j] = h->mu[j][0];
        out[5 * j + 1] = h->mu[j][1];
        out[5 * j + 2] = h->mu[j][2];
        out[5 * j + 3] = h->kappa[j];
        out[5 * j + 4] = h->pi[j];
    }
    out[40] = h->alpha;
    out[41] = out[42] = out[43] = 0.0;
}

void or_guide_forward(const double *params, const float bounds[6], const or_guide_sample *s, int32_t n,
                      int32_t emulate_bf16, double *raw, double *head)
{
    acts_t a;
    for (int32_t i = 0; i < n; ++i) {
        forward_one(params, bounds, s + i, emulate_bf16, &a);
        if (raw) memcpy(raw + 48 * (size_t)i, a.y, sizeof a.y);
        if (head) {
            head_t h;
            decode(a.y, &h);
            decode_out(&h, head + 44 * (size_t)i);
        }
    }
}

static double sample_pdf(const head_t *h, const or_guide_sample *s, double *q_out, double *b_out, double f[8])
{
    const double w[3] = {s->wi[0], s->wi[1], s->wi[2]};
    const double q = mixture(h, w, f);
    const double c = s->normal[0] * w[0] + s->normal[1] * w[1] + s->normal[2] * w[2];
    const double b = c > 0.0 ? c / ORPI : 0.0;
    *q_out = q;
    *b_out = b;
    return h->alpha * q + (1.0 - h->alpha) * b;
}

double or_guide_loss(const double *params, const double *frozen, const float bounds[6], const or_guide_sample *s,
                     int32_t n, int32_t emulate_bf16)
{
    return or_guide_loss_sel(params, frozen, bounds, s, n, emulate_bf16, 0);
}

double or_guide_loss_sel(const double *params, const double *frozen, const float bounds[6],
                         const or_guide_sample *s, int32_t n, int32_t emulate_bf16, int32_t objective)
{
    acts_t a, a0;
    double loss = 0.0;
    for (int32_t i = 0; i < n; ++i) {
        if (s[i].weight == 0.0f) continue;
        forward_one(params, bounds, s + i, emulate_bf16, &a);
        head_t h;
        decode(a.y, &h);
        double q, b, f[8];
        sample_pdf(&h, s + i, &q, &b, f);
        double q0 = q;
        if (frozen) {
            forward_one(frozen, bounds, s + i, emulate_bf16, &a0);
            head_t h0;
            decode(a0.y, &h0);
            double b0, f0[8];
            sample_pdf(&h0, s + i, &q0, &b0, f0);
        }
        const double w = (double)s[i].weight;
        loss += w * -log(Q_MIX * q + (1.0 - Q_MIX) * b);
        if (objective == 0) loss += w * -log(h.alpha * q0 + (1.0 - h.alpha) * b);
        if (objective == 1) loss += w * w * (double)s[i].pdf / (h.alpha * q0 + (1.0 - h.alpha) * b);
    }
    return loss;
}

/* d loss_i / d y (48 raw outputs) for one sample: the lobes get the gradient
 * of the fixed-proportion term, the selection logit that of the learned
 * one-sample MIS pdf with the mixture held fixed. */
static void head_backward(const head_t *h, const or_guide_sample *s, int objective, double *dy)
{
    for (int k = 0; k < 48; ++k) dy[k] = 0.0;
    const double wgt = s->weight;
    if (wgt == 0.0) return;
    double q, b, f[8];
    const double p = sample_pdf(h, s, &q, &b, f);
    const double G = -wgt / p;
    const double dq = -wgt / (Q_MIX * q + (1.0 - Q_MIX) * b) * Q_MIX;
    /* second-moment objective: d/d alpha of w^2 ps / p is the KL's -w (q - b) / p
     * times w ps / p */
    const double rs = objective == 1 ? (p > 0.0 ? wgt * (double)s->pdf / p : 0.0) : 1.0;
    dy[SEL_ROW] = G * rs * (q - b) * A_SPAN * h->sig * (1.0 - h->sig);
    const double w[3] = {s->wi[0], s->wi[1], s->wi[2]};
    for (int j = 0; j < 8; ++j) {
        const double t = h->pi[j] * f[j];
        const double kap = h->kappa[j];
        const double c = h->mu[j][0] * w[0] + h->mu[j][1] * w[1] + h->mu[j][2] * w[2];
        dy[head_row(j, 4)] = dq * (t - h->pi[j] * q);
        if (h->kr[j] > KR_LO && h->kr[j] < KR_HI)
            dy[head_row(j, 3)] = dq * t * (1.0 - 2.0 * kap / expm1(2.0 * kap) + kap * (c - 1.0));
        double g[3];
        for (int k = 0; k < 3; ++k) g[k] = dq * t * kap * w[k];
        const double mg = h->mu[j][0] * g[0] + h->mu[j][1] * g[1] + h->mu[j][2] * g[2];
        for (int k = 0; k < 3; ++k) dy[head_row(j, k)] = (g[k] - h->mu[j][k] * mg) / h->r[j];
    }
}

double or_guide_grad(const double *params, const float bounds[6], const or_guide_sample *s, int32_t n,
                     int32_t emulate_bf16, double *grad)
{
    return or_guide_grad_sel(params, bounds, s, n, emulate_bf16, 0, grad);
}

double or_guide_grad_sel(const double *params, const float bounds[6], const or_guide_sample *s, int32_t n,
                         int32_t emulate_bf16, int32_t objective, double *grad)
{
    const int emu = emulate_bf16;
    memset(grad, 0, sizeof(double) * OR_GUIDE_NP);
    double sumw = 0.0;
    acts_t a;
    double dcur[64], dnext[64];
    for (int32_t i = 0; i < n; ++i) {
        sumw += (double)s[i].weight;
        if (s[i].weight == 0.0f) continue;
        forward_one(params, bounds, s + i, emu, &a);
        head_t h;
        decode(a.y, &h);
        head_backward(&h, s + i, objective, dcur);
        if (emu)
            for (int k = 0; k < 48; ++k) dcur[k] = bf16r(dcur[k]);
        /* dcur = d loss / d z_{l+1} (pre-activation of layer l's output) */
        for (int l = NL - 1; l >= 0; --l) {
            const double *W = params + w_off(l);
            double *gW = grad + w_off(l);
            double *gb = grad + b_off(l);
            const int ni = G_IN[l], no = G_OUT[l];
            for (int o = 0; o < no; ++o) {
                gb[o] += dcur[o];
                for (int k = 0; k < ni; ++k) gW[o * ni + k] += dcur[o] * a.x[l][k];
            }
            if (l == 0) break;
            /* the second-moment objective trains the selection row alone */
            if (l == NL - 1 && objective == 1) dcur[SEL_ROW] = 0.0;
            for (int k = 0; k < ni; ++k) {
                double d = 0.0;
                for (int o = 0; o < no; ++o) d += (emu ? bf16r(W[o * ni + k]) : W[o * ni + k]) * dcur[o];
                d = a.x[l][k] > 0.0 ? d : 0.0;
                dnext[k] = emu ? bf16r(d) : d;
            }
            memcpy(dcur, dnext, sizeof(double) * (size_t)ni);
        }
    }
    return sumw;
}
