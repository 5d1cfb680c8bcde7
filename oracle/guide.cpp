// ORACLE — test infrastructure only (see pbrt.h). The learned sampling
// proposal (not in the reference): frequency encoding of the hit position and
// normal, a 144-64-64-64-40 ReLU MLP, an 8-lobe von Mises-Fisher mixture on
// the sphere of directions, its weighted negative-log-likelihood loss with
// analytic gradients, and Adam (Kingma & Ba 2015). Templated on the scalar
// type so the same code runs in f64 (finite-difference gradient checks) and
// in f32 with optional bf16 rounding of weights and layer inputs at the
// points where the GPU kernels round them.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int kIn = 6, kOct = 12, kEnc = kIn * kOct * 2;  // 144
constexpr int kHid = 64, kLobes = 8, kOut = 5 * kLobes;   // 40
constexpr int kLayers = 4;
const int kDims[kLayers + 1] = {kEnc, kHid, kHid, kHid, kOut};
// parameter offsets: W_l is [in][out] row-major, then b_l [out]
int w_off(int l) {
    int o = 0;
    for (int i = 0; i < l; ++i) o += kDims[i] * kDims[i + 1] + kDims[i + 1];
    return o;
}
int b_off(int l) { return w_off(l) + kDims[l] * kDims[l + 1]; }
constexpr int kParams = 144 * 64 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 * 40 + 40;  // 20200

float bf16_round(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;  // inf / nan unchanged
    uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;  // round to nearest even
    u &= 0xffff0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

template <typename T>
T rnd(T x, bool bf16) {
    return bf16 ? (T)bf16_round((float)x) : x;
}

// Position normalised to [0,1]^3 by the scene bounds, normal in [-1,1]^3;
// feature (6*i + ... ) order: input i, octave j -> sin(2^j pi v), cos(2^j pi v)
// at index i*24 + 2j, i*24 + 2j + 1. The angle is reduced exactly first:
// 2^j pi v = 2 pi (2^(j-1) v), and 2^(j-1) v and its fractional part are
// exact in binary floating point, so sin/cos take 2 pi fract(2^(j-1) v).
template <typename T>
void encode(const float *x, const float *n, const float *bounds, T *f) {
    T v[kIn];
    for (int a = 0; a < 3; ++a) {
        float ext = bounds[3 + a] - bounds[a];
        v[a] = (T)((x[a] - bounds[a]) / (ext > 0 ? ext : 1.f));
        v[3 + a] = (T)n[a];
    }
    for (int i = 0; i < kIn; ++i)
        for (int j = 0; j < kOct; ++j) {
            const double r = std::ldexp((double)v[i], j - 1);
            const double arg = 6.28318530717958647692 * (r - std::floor(r));
            f[i * 2 * kOct + 2 * j] = (T)std::sin(arg);
            f[i * 2 * kOct + 2 * j + 1] = (T)std::cos(arg);
        }
}

template <typename T>
struct Fwd {
    T a[kLayers + 1][kEnc];  // layer inputs (post-activation), a[0] = encoding
    T z[kLayers][kHid];      // pre-activations
    T out[kOut];
};

template <typename T>
void forward(const T *p, const float *x, const float *n, const float *bounds, bool bf16, Fwd<T> &F) {
    encode<T>(x, n, bounds, F.a[0]);
    for (int i = 0; i < kEnc; ++i) F.a[0][i] = rnd(F.a[0][i], bf16);
    for (int l = 0; l < kLayers; ++l) {
        int din = kDims[l], dout = kDims[l + 1];
        const T *W = p + w_off(l), *b = p + b_off(l);
        for (int o = 0; o < dout; ++o) {
            T s = 0;
            for (int i = 0; i < din; ++i) s += F.a[l][i] * rnd(W[i * dout + o], bf16);
            s += b[o];
            if (l < kLayers - 1) {
                F.z[l][o] = s;
                F.a[l + 1][o] = rnd(s > 0 ? s : (T)0, bf16);
            } else {
                F.out[o] = s;
            }
        }
    }
}

// Mixture parameters from the 40 outputs: per lobe j, out[5j] logit,
// out[5j+1..3] unnormalised mean direction, out[5j+4] log concentration
// (clamped to [-4.6, 6.9], kappa in [0.01, 1000]).
template <typename T>
struct Mix {
    T pi[kLobes], mu[kLobes][3], kappa[kLobes], mnorm[kLobes], lk[kLobes];
    bool clamped[kLobes];
};
template <typename T>
void decode(const T *out, Mix<T> &m) {
    T mx = out[0];
    for (int j = 1; j < kLobes; ++j) mx = std::max(mx, out[5 * j]);
    T s = 0;
    for (int j = 0; j < kLobes; ++j) {
        m.pi[j] = std::exp(out[5 * j] - mx);
        s += m.pi[j];
    }
    for (int j = 0; j < kLobes; ++j) {
        m.pi[j] /= s;
        T a = out[5 * j + 1], b = out[5 * j + 2], c = out[5 * j + 3];
        T len = std::sqrt(a * a + b * b + c * c);
        m.mnorm[j] = len;
        T inv = len > 0 ? 1 / len : 0;
        m.mu[j][0] = a * inv;
        m.mu[j][1] = b * inv;
        m.mu[j][2] = len > 0 ? c * inv : (T)1;
        T r = out[5 * j + 4];
        m.clamped[j] = r < (T)-4.6 || r > (T)6.9;
        m.lk[j] = std::min(std::max(r, (T)-4.6), (T)6.9);
        m.kappa[j] = std::exp(m.lk[j]);
    }
}
// vMF density on S^2: kappa / (2 pi (1 - e^{-2 kappa})) * exp(kappa (mu.w - 1))
template <typename T>
T vmf_pdf(const T *mu, T kappa, const float *w) {
    T c = mu[0] * (T)w[0] + mu[1] * (T)w[1] + mu[2] * (T)w[2];
    return kappa / ((T)(2 * 3.14159265358979323846) * (-std::expm1(-2 * kappa))) * std::exp(kappa * (c - 1));
}
template <typename T>
T mix_pdf(const Mix<T> &m, const float *w) {
    T q = 0;
    for (int j = 0; j < kLobes; ++j) q += m.pi[j] * vmf_pdf(m.mu[j], m.kappa[j], w);
    return q;
}

// The mixture times a vMF lobe (mu_b, k_b) fitted to a vertex's BSDF: lobe j
// becomes the vMF with k mu = k_j mu_j + k_b mu_b, its weight scaled by the
// integral of the two densities' product, c(k_j) c(k_b) exp(k - k_j - k_b) /
// c(k) with c(k) = k / (2 pi (1 - e^{-2k})), then renormalised. False if the
// weights degenerate.
template <typename T>
T vmf_norm(T k) {
    return k / ((T)(2 * 3.14159265358979323846) * (-std::expm1(-2 * k)));
}
template <typename T>
bool product(Mix<T> &m, const float *mu_b, T k_b) {
    T s = 0;
    const T cb = vmf_norm(k_b);
    for (int j = 0; j < kLobes; ++j) {
        const T kj = m.kappa[j];
        const T vx = kj * m.mu[j][0] + k_b * (T)mu_b[0], vy = kj * m.mu[j][1] + k_b * (T)mu_b[1],
                vz = kj * m.mu[j][2] + k_b * (T)mu_b[2];
        const T k = std::max(std::sqrt(vx * vx + vy * vy + vz * vz), (T)1e-4);
        const T scale = vmf_norm(kj) * cb * std::exp(std::min(k - kj - k_b, (T)0)) / vmf_norm(k);
        m.mu[j][0] = vx / k;
        m.mu[j][1] = vy / k;
        m.mu[j][2] = vz / k;
        m.kappa[j] = k;
        m.pi[j] *= scale;
        s += m.pi[j];
    }
    if (!(s > 0) || !std::isfinite(s)) return false;
    const T inv = 1 / s;
    for (int j = 0; j < kLobes; ++j) m.pi[j] *= inv;
    return true;
}

// d(-log q)/d out, for one direction w.
template <typename T>
void nll_grad(const T *out, const Mix<T> &m, const float *w, T *g) {
    T v[kLobes], q = 0;
    for (int j = 0; j < kLobes; ++j) {
        v[j] = vmf_pdf(m.mu[j], m.kappa[j], w);
        q += m.pi[j] * v[j];
    }
    for (int j = 0; j < kLobes; ++j) {
        T gam = m.pi[j] * v[j] / q;
        g[5 * j] = m.pi[j] - gam;
        T k = m.kappa[j];
        T c = m.mu[j][0] * (T)w[0] + m.mu[j][1] * (T)w[1] + m.mu[j][2] * (T)w[2];
        // d log v / d kappa = 1/kappa - 2 e^{-2k}/(1 - e^{-2k}) + (c - 1)
        T e2 = std::exp(-2 * k);
        T dlk = 1 / k - 2 * e2 / (-std::expm1(-2 * k)) + (c - 1);
        g[5 * j + 4] = m.clamped[j] ? (T)0 : -gam * dlk * k;
        // d log v / d mu = kappa * w, projected through mu = m/|m|
        T inv = m.mnorm[j] > 0 ? 1 / m.mnorm[j] : 0;
        for (int a = 0; a < 3; ++a) {
            T proj = (T)w[a] - c * m.mu[j][a];
            g[5 * j + 1 + a] = -gam * k * proj * inv;
        }
    }
    (void)out;
}

template <typename T>
void backward(const T *p, const Fwd<T> &F, const T *gout, bool bf16, T *grad) {
    T g[kHid > kOut ? kHid : kOut];
    for (int o = 0; o < kOut; ++o) g[o] = gout[o];
    for (int l = kLayers - 1; l >= 0; --l) {
        int din = kDims[l], dout = kDims[l + 1];
        const T *W = p + w_off(l);
        T *gW = grad + w_off(l), *gb = grad + b_off(l);
        for (int o = 0; o < dout; ++o) gb[o] += g[o];
        for (int i = 0; i < din; ++i)
            for (int o = 0; o < dout; ++o) gW[i * dout + o] += F.a[l][i] * g[o];
        if (l == 0) break;
        T gi[kHid];
        for (int i = 0; i < din; ++i) {
            T s = 0;
            for (int o = 0; o < dout; ++o) s += rnd(W[i * dout + o], bf16) * g[o];
            gi[i] = F.z[l - 1][i] > 0 ? s : (T)0;
        }
        for (int i = 0; i < din; ++i) g[i] = gi[i];
    }
}

template <typename T>
T loss_and_grad(const T *p, int n, const float *x, const float *nrm, const float *wi, const float *weight,
                const float *bounds, bool bf16, T *grad) {
    T wsum = 0;
    for (int r = 0; r < n; ++r) wsum += (T)weight[r];
    if (!(wsum > 0)) return 0;
    if (grad) std::memset(grad, 0, sizeof(T) * kParams);
    T loss = 0;
    for (int r = 0; r < n; ++r) {
        if (!(weight[r] > 0)) continue;
        Fwd<T> F;
        forward<T>(p, x + 3 * r, nrm + 3 * r, bounds, bf16, F);
        Mix<T> m;
        decode<T>(F.out, m);
        T s = (T)weight[r] / wsum;
        loss += -s * std::log(mix_pdf(m, wi + 3 * r));
        if (grad) {
            T g[kOut];
            nll_grad<T>(F.out, m, wi + 3 * r, g);
            for (int o = 0; o < kOut; ++o) g[o] *= s;
            backward<T>(p, F, g, bf16, grad);
        }
    }
    return loss;
}

void coordinate_system(const float *v1, float *v2, float *v3) {
    if (std::fabs(v1[0]) > std::fabs(v1[1])) {
        float l = std::sqrt(v1[0] * v1[0] + v1[2] * v1[2]);
        v2[0] = -v1[2] / l; v2[1] = 0; v2[2] = v1[0] / l;
    } else {
        float l = std::sqrt(v1[1] * v1[1] + v1[2] * v1[2]);
        v2[0] = 0; v2[1] = v1[2] / l; v2[2] = -v1[1] / l;
    }
    v3[0] = v1[1] * v2[2] - v1[2] * v2[1];
    v3[1] = v1[2] * v2[0] - v1[0] * v2[2];
    v3[2] = v1[0] * v2[1] - v1[1] * v2[0];
}

}  // namespace

extern "C" {

int orc_guide_num_params() { return kParams; }

void orc_guide_forward(const float *params, int n, const float *x, const float *nrm, const float *bounds, int bf16,
                       float *out) {
    for (int r = 0; r < n; ++r) {
        Fwd<float> F;
        forward<float>(params, x + 3 * r, nrm + 3 * r, bounds, bf16 != 0, F);
        std::memcpy(out + kOut * r, F.out, sizeof(float) * kOut);
    }
}

void orc_guide_encode(int n, const float *x, const float *nrm, const float *bounds, float *feat) {
    for (int r = 0; r < n; ++r) encode<float>(x + 3 * r, nrm + 3 * r, bounds, feat + kEnc * r);
}

// Weighted NLL: loss = -sum_r (w_r / sum w) log q(wi_r | x_r, n_r). Returns the
// loss; grad (kParams) receives its gradient. f64 != 0 computes in double.
double orc_guide_loss_grad(const float *params, int n, const float *x, const float *nrm, const float *wi,
                           const float *weight, const float *bounds, int bf16, int f64, float *grad) {
    if (f64) {
        std::vector<double> p(params, params + kParams), g(kParams, 0.0);
        double l = loss_and_grad<double>(p.data(), n, x, nrm, wi, weight, bounds, false, grad ? g.data() : nullptr);
        if (grad)
            for (int i = 0; i < kParams; ++i) grad[i] = (float)g[i];
        return l;
    }
    return loss_and_grad<float>(params, n, x, nrm, wi, weight, bounds, bf16 != 0, grad);
}

// Same loss with the parameters given in f64 (finite differences).
double orc_guide_loss_f64(const double *params, int n, const float *x, const float *nrm, const float *wi,
                          const float *weight, const float *bounds) {
    return loss_and_grad<double>(params, n, x, nrm, wi, weight, bounds, false, nullptr);
}

// Mixture pdf of directions given raw network outputs (n x 40).
void orc_guide_pdf(const float *outs, int n, const float *wi, float *pdf) {
    for (int r = 0; r < n; ++r) {
        Mix<double> m;
        double o[kOut];
        for (int k = 0; k < kOut; ++k) o[k] = outs[kOut * r + k];
        decode<double>(o, m);
        pdf[r] = (float)mix_pdf(m, wi + 3 * r);
    }
}

// Pdf of the mixture times the vMF lobe (lobe_mu[3r..], lobe_k[r]) per row;
// NaN where the product degenerates.
void orc_guide_product_pdf(const float *outs, int n, const float *lobe_mu, const float *lobe_k, const float *wi,
                           float *pdf) {
    for (int r = 0; r < n; ++r) {
        Mix<double> m;
        double o[kOut];
        for (int k = 0; k < kOut; ++k) o[k] = outs[kOut * r + k];
        decode<double>(o, m);
        pdf[r] = product<double>(m, lobe_mu + 3 * r, (double)lobe_k[r]) ? (float)mix_pdf(m, wi + 3 * r) : NAN;
    }
}

static void sample_rows(const float *outs, int n, const float *u, float *wi, float *pdf, const float *lobe_mu,
                        const float *lobe_k);

// Sample the mixture: u = (lobe, u1, u2) per row -> direction and its mixture pdf.
void orc_guide_sample(const float *outs, int n, const float *u, float *wi, float *pdf) {
    sample_rows(outs, n, u, wi, pdf, nullptr, nullptr);
}
// Sample the mixture times the vMF lobe of each row, with the product's pdf.
void orc_guide_product_sample(const float *outs, int n, const float *lobe_mu, const float *lobe_k, const float *u,
                              float *wi, float *pdf) {
    sample_rows(outs, n, u, wi, pdf, lobe_mu, lobe_k);
}

}  // extern "C"

static void sample_rows(const float *outs, int n, const float *u, float *wi, float *pdf, const float *lobe_mu,
                        const float *lobe_k) {
    for (int r = 0; r < n; ++r) {
        Mix<float> m;
        decode<float>(outs + kOut * r, m);
        if (lobe_mu && !product<float>(m, lobe_mu + 3 * r, lobe_k[r])) {
            wi[3 * r] = wi[3 * r + 1] = wi[3 * r + 2] = 0.f;
            pdf[r] = NAN;
            continue;
        }
        float ul = u[3 * r], u1 = u[3 * r + 1], u2 = u[3 * r + 2];
        int j = 0;
        float c = 0;
        for (; j < kLobes - 1; ++j) {
            c += m.pi[j];
            if (ul < c) break;
        }
        float k = m.kappa[j];
        // Jakob 2012: w = 1 + log(u1 + (1 - u1) e^{-2 kappa}) / kappa
        float w = 1 + std::log(u1 + (1 - u1) * std::exp(-2 * k)) / k;
        w = std::fmin(1.f, std::fmax(-1.f, w));
        float s = std::sqrt(std::fmax(0.f, 1 - w * w));
        float phi = 2 * 3.14159265358979323846f * u2;
        float t1[3], t2[3];
        coordinate_system(m.mu[j], t1, t2);
        float lx = s * std::cos(phi), ly = s * std::sin(phi);
        for (int a = 0; a < 3; ++a) wi[3 * r + a] = t1[a] * lx + t2[a] * ly + m.mu[j][a] * w;
        pdf[r] = (float)mix_pdf(m, wi + 3 * r);
    }
}

extern "C" {

// Adam (Kingma & Ba 2015), bias-corrected, step t >= 1.
void orc_adam(float *p, float *m, float *v, const float *g, int n, int t, float lr, float b1, float b2, float eps) {
    float c1 = 1 - std::pow(b1, (float)t), c2 = 1 - std::pow(b2, (float)t);
    for (int i = 0; i < n; ++i) {
        m[i] = b1 * m[i] + (1 - b1) * g[i];
        v[i] = b2 * v[i] + (1 - b2) * g[i] * g[i];
        float mh = m[i] / c1, vh = v[i] / c2;
        p[i] -= lr * mh / (std::sqrt(vh) + eps);
    }
}

}  // extern "C"
