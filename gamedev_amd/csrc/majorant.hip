// Learned rate majorant: training and inference kernels of the 8-64-64-1 perceptron (gdx_mlp.h) and the
// C ABI of a majorant model (gdx_col_majorant_*). No counterpart in the reference.
//
//   k_mlp_infer   y = softplus(net(x)) for n feature rows, one 16-row tile per wavefront step
//   k_mlp_grad    forward + backward of a minibatch: each wavefront takes TILES_PER_WAVE tiles and writes
//                 one partial gradient row (deterministic, no atomics); weight gradients are MFMA
//                 products over the sample dimension, their operands transposed through LDS
//   k_mlp_reduce  sum of the partial rows in wavefront order (gradient check)
//   k_mlp_adam    the same sum, then Adam (Kingma and Ba 2015, bias-corrected) on the FP32 master
//                 parameters and the BF16 fragments rebuilt, in one block
// Loss of a sample: (ln softplus(z) - ln t)^2, the mean over the minibatch.
#include "gdx_internal.h"
#include "gdx_mlp.h"

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

// Scratch of gdx_col_majorant_grad_noalloc: one partial gradient row per wavefront for up to
// GDX_COL_MAJORANT_GRAD_NOALLOC_MAX samples (128 samples per wavefront), then the reduced row (gradient
// and loss sum). Every row the kernels read is written first by the same call, so nothing is cleared.
constexpr int GRAD_NOALLOC_ROWS = GDX_COL_MAJORANT_GRAD_NOALLOC_MAX / 128;
constexpr size_t GRAD_NOALLOC_FLOATS = size_t(GRAD_NOALLOC_ROWS + 1)*gdx::mlp::P_STRIDE;
static __device__ __align__(256) float scratch[GRAD_NOALLOC_FLOATS];

namespace gdx {
namespace mlp {
namespace {

constexpr int BLOCK = 256;
constexpr int WAVES = BLOCK / 64;
constexpr int TILES_PER_WAVE = 8;

__device__ void build_pack(const float *p, const float *mu, const float *inv_sigma, pack *P, int tid, int nthreads)
{
    for (int e = tid; e < 4*64; e += nthreads)
    {
        int c = e / 64, lane = e % 64, g = lane >> 4, i = lane & 15;
        v4s r;
        for (int v = 0; v < 4; ++v)
        {
            int k = 4*g + v;
            r[v] = k < F ? bf16_rn(p[P_W1 + k*H + 16*c + i]) : short(0);
        }
        P->w1a[c][lane] = r;
    }
    for (int e = tid; e < 16*64; e += nthreads)
    {
        int t = e / 64, lane = e % 64, g = lane >> 4, i = lane & 15;
        int hi = t / 4, lo = t % 4;
        v4s a, b;
        for (int v = 0; v < 4; ++v)
        {
            a[v] = bf16_rn(p[P_W2 + (16*lo + 4*g + v)*H + 16*hi + i]);   // m = hi, c = lo
            b[v] = bf16_rn(p[P_W2 + (16*hi + i)*H + 16*lo + 4*g + v]);   // a = hi, b = lo
        }
        P->w2a[t][lane] = a;
        P->w2b[t][lane] = b;
    }
    for (int e = tid; e < H; e += nthreads)
    {
        P->b1[e] = p[P_B1 + e];
        P->b2[e] = p[P_B2 + e];
        P->w3[e] = p[P_W3 + e];
    }
    if (tid == 0) P->b3 = p[P_B3];
    if (tid < F) { P->mu[tid] = mu[tid]; P->inv_sigma[tid] = inv_sigma[tid]; }
}

__global__ void k_mlp_pack(const float *p, const float *norm, pack *P)
{
    build_pack(p, norm, norm + F, P, threadIdx.x, blockDim.x);
}

__global__ __launch_bounds__(BLOCK) void k_mlp_infer(const pack *P, const float *X, long long n, float *y)
{
    const int lane = threadIdx.x & 63, g = lane >> 4, j = lane & 15;
    const long long tiles = (n + 15) / 16;
    for (long long t = (long long)blockIdx.x*WAVES + threadIdx.x/64; t < tiles; t += (long long)gridDim.x*WAVES)
    {
        long long r = t*16 + j;
        bool valid = r < n;
        v4s x = input_fragment(P, X + (valid ? r : 0)*F, valid);
        forward_state f;
        float z = tile_forward(P, x, f);
        if (g == 0 && valid) y[r] = softplus(z);
    }
}

// sum over the 16 lanes of a row (lanes sharing lane / 16)
__device__ __forceinline__ float row_sum(float v)
{
    v += __shfl_xor(v, 1, 64);
    v += __shfl_xor(v, 2, 64);
    v += __shfl_xor(v, 4, 64);
    v += __shfl_xor(v, 8, 64);
    return v;
}

// forward + backward over the samples idx[start .. start + count) (idx null: identity); one partial row
// per wavefront of the grid. lt holds ln(target).
__global__ __launch_bounds__(BLOCK) void k_mlp_grad(const pack *P, const float *X, const float *lt, const int *idx,
                                                   long long start, int count, float *partial)
{
    __shared__ short sH1[WAVES][16][H], sDA2[WAVES][16][H], sDA1[WAVES][16][H], sX[WAVES][16][F];
    const int lane = threadIdx.x & 63, g = lane >> 4, j = lane & 15, wv = threadIdx.x / 64;
    const int wave = blockIdx.x*WAVES + wv;
    const float inv_count = 1.0f / float(count);

    v4f acc2[4][4], acc1[4];
    float gw3[4][4], gb2[4][4], gb1[4][4], gb3 = 0.0f, loss = 0.0f;
#pragma unroll
    for (int a = 0; a < 4; ++a)
    {
        acc1[a] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int b = 0; b < 4; ++b)
        {
            acc2[a][b] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
            gw3[a][b] = gb2[a][b] = gb1[a][b] = 0.0f;
        }
    }

    for (int tt = 0; tt < TILES_PER_WAVE; ++tt)
    {
        const int row = (wave*TILES_PER_WAVE + tt)*16 + j;
        const bool valid = row < count;
        long long r = 0;
        if (valid) r = idx ? (long long)idx[start + row] : start + row;
        v4s x = input_fragment(P, X + r*F, valid);
        forward_state f;
        float z = tile_forward(P, x, f);
        float y = softplus(z);
        float d = valid ? logf(y) - lt[r] : 0.0f;
        float dz = valid ? 2.0f*d*sigmoid(z)/y*inv_count : 0.0f;
        if (g == 0) { loss += d*d; gb3 += dz; }

        // layer 3 and the layer-2 pre-activation gradient (C layout)
        v4s da2b[4];
#pragma unroll
        for (int m = 0; m < 4; ++m)
        {
            v4s q;
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                float a2 = f.a2[m][v];
                gw3[m][v] += dz*silu(a2);
                float da2 = dz*P->w3[16*m + 4*g + v]*silu_grad(a2);
                gb2[m][v] += da2;
                q[v] = bf16_rn(da2);
            }
            da2b[m] = q;
        }
        // dH1^T = W2 dA2^T, then the layer-1 pre-activation gradient
        v4s da1b[4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
        {
            v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int b = 0; b < 4; ++b) acc = mfma(P->w2b[4*a + b][lane], da2b[b], acc);
            v4s q;
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                float da1 = acc[v]*silu_grad(f.a1[a][v]);
                gb1[a][v] += da1;
                q[v] = bf16_rn(da1);
            }
            da1b[a] = q;
        }
        // weight gradients contract over the tile's samples: operands move through LDS to the
        // sample-as-k layout
#pragma unroll
        for (int a = 0; a < 4; ++a)
        {
            *reinterpret_cast<v4s *>(&sH1[wv][j][16*a + 4*g]) = f.h1[a];
            *reinterpret_cast<v4s *>(&sDA2[wv][j][16*a + 4*g]) = da2b[a];
            *reinterpret_cast<v4s *>(&sDA1[wv][j][16*a + 4*g]) = da1b[a];
        }
        if (g < F/4) *reinterpret_cast<v4s *>(&sX[wv][j][4*g]) = x;
        __syncthreads();
        v4s xa = {0, 0, 0, 0};
#pragma unroll
        for (int v = 0; v < 4; ++v) if (j < F) xa[v] = sX[wv][4*g + v][j];
#pragma unroll
        for (int a = 0; a < 4; ++a)
        {
            v4s ha, d1;
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                ha[v] = sH1[wv][4*g + v][16*a + j];
                d1[v] = sDA1[wv][4*g + v][16*a + j];
            }
            // dW1[k][16a + n] (k = input), dW2 rows 16a.. against every column tile b
            acc1[a] = mfma(xa, d1, acc1[a]);
#pragma unroll
            for (int b = 0; b < 4; ++b)
            {
                v4s dbb;
#pragma unroll
                for (int v = 0; v < 4; ++v) dbb[v] = sDA2[wv][4*g + v][16*b + j];
                acc2[a][b] = mfma(ha, dbb, acc2[a][b]);
            }
        }
        __syncthreads();
    }

    float *out = partial + (long long)wave*P_STRIDE;
#pragma unroll
    for (int a = 0; a < 4; ++a)
    {
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            int hrow = 16*a + 4*g + v;
#pragma unroll
            for (int b = 0; b < 4; ++b) out[P_W2 + hrow*H + 16*b + j] = acc2[a][b][v];
            if (g < F/4) out[P_W1 + (4*g + v)*H + 16*a + j] = acc1[a][v];
            float s3 = row_sum(gw3[a][v]), s2 = row_sum(gb2[a][v]), s1 = row_sum(gb1[a][v]);
            if (j == 0)
            {
                out[P_W3 + hrow] = s3;
                out[P_B2 + hrow] = s2;
                out[P_B1 + hrow] = s1;
            }
        }
    }
    gb3 = row_sum(gb3);
    loss = row_sum(loss);
    if (lane == 0) { out[P_B3] = gb3; out[P_LOSS] = loss; }
}

// gradient and loss sum of one minibatch: partial rows summed in wavefront order
__global__ void k_mlp_reduce(const float *partial, int rows, float *grad)
{
    for (int p = blockIdx.x*blockDim.x + threadIdx.x; p <= P_LOSS; p += gridDim.x*blockDim.x)
    {
        float s = 0.0f;
        for (int w = 0; w < rows; ++w) s += partial[(long long)w*P_STRIDE + p];
        grad[p] = s;
    }
}

struct adam_args
{
    float lr, beta1, beta2, eps, bc1, bc2;
};

__global__ __launch_bounds__(1024) void k_mlp_adam(float *p, float *m1, float *m2, const float *partial, int rows,
                                                  adam_args A, const float *norm, pack *P, float *loss_out)
{
    for (int e = threadIdx.x; e <= P_LOSS; e += blockDim.x)
    {
        float g = 0.0f;
        for (int w = 0; w < rows; ++w) g += partial[(long long)w*P_STRIDE + e];
        if (e == P_LOSS)
        {
            if (loss_out) *loss_out = g;
            continue;
        }
        float m = A.beta1*m1[e] + (1.0f - A.beta1)*g;
        float v = A.beta2*m2[e] + (1.0f - A.beta2)*(g*g);
        m1[e] = m;
        m2[e] = v;
        p[e] = p[e] - A.lr*(m / A.bc1) / (sqrtf(v / A.bc2) + A.eps);
    }
    __syncthreads();
    build_pack(p, norm, norm + F, P, threadIdx.x, blockDim.x);
}

uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30))*0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27))*0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
double unit(uint64_t &s) { return double(splitmix(s) >> 11)*0x1.0p-53; }

}  // namespace

// one majorant model: FP32 parameters and Adam moments, the packed fragments, the input normalization
struct model
{
    dbuf<float> p, m1, m2, norm, partial, grad, loss;
    dbuf<pack> P;
    dbuf<int> perm;
    long long steps = 0;
    unsigned long long version = 0;     // advanced by every change of the parameters or normalization
    std::vector<float> h_norm = std::vector<float>(2*F);
    std::string last_error;
};

int grid_waves(int count) { return (count + 16*TILES_PER_WAVE - 1) / (16*TILES_PER_WAVE); }

void model_init(model &M, uint64_t seed)
{
    // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases; the output bias puts softplus at 1
    std::vector<float> h(P_COUNT);
    uint64_t s = seed;
    auto fill = [&](int off, int count, int fan_in) {
        double bound = 1.0 / std::sqrt(double(fan_in));
        for (int i = 0; i < count; ++i) h[off + i] = float((2.0*unit(s) - 1.0)*bound);
    };
    fill(P_W1, F*H, F);
    fill(P_B1, H, F);
    fill(P_W2, H*H, H);
    fill(P_B2, H, H);
    fill(P_W3, H, H);
    h[P_B3] = 0.54132485f;       // softplus^-1(1)
    M.p.alloc(P_COUNT);
    M.m1.alloc(P_COUNT);
    M.m2.alloc(P_COUNT);
    M.norm.alloc(2*F);
    M.P.alloc(1);
    M.grad.alloc(P_COUNT + 1);
    for (int k = 0; k < F; ++k) { M.h_norm[k] = 0.0f; M.h_norm[F + k] = 1.0f; }
    hip_check(hipMemcpy(M.p.p, h.data(), sizeof(float)*P_COUNT, hipMemcpyHostToDevice), "majorant init");
    hip_check(hipMemset(M.m1.p, 0, sizeof(float)*P_COUNT), "majorant init");
    hip_check(hipMemset(M.m2.p, 0, sizeof(float)*P_COUNT), "majorant init");
    hip_check(hipMemcpy(M.norm.p, M.h_norm.data(), sizeof(float)*2*F, hipMemcpyHostToDevice), "majorant init");
    k_mlp_pack<<<1, 256>>>(M.p.p, M.norm.p, M.P.p);
    hip_check(hipGetLastError(), "majorant pack");
    hip_check(hipDeviceSynchronize(), "majorant pack");
}

const pack *model_pack(const model &M) { return M.P.p; }

void launch_mlp_infer(const pack *P, const float *X, long long n, float *y, hipStream_t st)
{
    if (n <= 0) return;
    long long tiles = (n + 15) / 16;
    long long blocks = std::min<long long>((tiles + WAVES - 1) / WAVES, 8192);
    k_mlp_infer<<<dim3(unsigned(blocks)), BLOCK, 0, st>>>(P, X, n, y);
    hip_check(hipGetLastError(), "majorant inference launch");
}

}  // namespace mlp
}  // namespace gdx

using gdx::mlp::model;

struct gdx_col_majorant_ctx
{
    model m;
};

namespace {

thread_local std::string g_majorant_error;

template <typename F>
int mguard(gdx_col_majorant_ctx *h, F &&f)
{
    try
    {
        f();
        return GDX_COL_OK;
    }
    catch (const gdx::status_error &e)
    {
        (h ? h->m.last_error : g_majorant_error) = e.what();
        return e.code;
    }
    catch (const std::bad_alloc &)
    {
        (h ? h->m.last_error : g_majorant_error) = "host allocation failure";
        return GDX_COL_ERR_ALLOC;
    }
    catch (const std::exception &e)
    {
        (h ? h->m.last_error : g_majorant_error) = e.what();
        return GDX_COL_ERR_HIP;
    }
}

void check_arg(bool ok, const char *what)
{
    if (!ok) throw gdx::status_error(GDX_COL_ERR_ARG, what);
}

// Device address of `scratch` on the current GPU. The array's name is a host-side stand-in in host code;
// every GPU holds its own copy at its own address, looked up once per device with that device current.
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[MAX_DEVICES] = {};

float *scratch_on_current_device()
{
    int dev = -1;
    gdx::hip_check(hipGetDevice(&dev), "current device");
    if (dev < 0 || dev >= MAX_DEVICES) throw gdx::status_error(GDX_COL_ERR_ARG, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev])
    {
        void *p = nullptr;
        gdx::hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch address");
        g_scratch_ptr[dev] = static_cast<float *>(p);
    }
    return g_scratch_ptr[dev];
}

}  // namespace

extern "C" {

int gdx_col_majorant_create(unsigned long long seed, gdx_col_majorant *out)
{
    if (!out) return GDX_COL_ERR_ARG;
    *out = nullptr;
    gdx_col_majorant_ctx *h = nullptr;
    int rc = mguard(nullptr, [&] {
        h = new gdx_col_majorant_ctx();
        gdx::mlp::model_init(h->m, seed);
    });
    if (rc != GDX_COL_OK) { delete h; return rc; }
    *out = h;
    return GDX_COL_OK;
}

int gdx_col_majorant_destroy(gdx_col_majorant h)
{
    delete h;
    return GDX_COL_OK;
}

const char *gdx_col_majorant_last_error(gdx_col_majorant h)
{
    return h ? h->m.last_error.c_str() : g_majorant_error.c_str();
}

int gdx_col_majorant_get_params(gdx_col_majorant h, float *params, float *norm)
{
    if (!h) return GDX_COL_ERR_ARG;
    return mguard(h, [&] {
        if (params)
            gdx::hip_check(hipMemcpy(params, h->m.p.p, sizeof(float)*gdx::mlp::P_COUNT, hipMemcpyDeviceToHost),
                           "majorant parameters");
        if (norm) std::memcpy(norm, h->m.h_norm.data(), sizeof(float)*2*gdx::mlp::F);
    });
}

int gdx_col_majorant_set_params(gdx_col_majorant h, const float *params, const float *norm)
{
    if (!h) return GDX_COL_ERR_ARG;
    return mguard(h, [&] {
        using namespace gdx::mlp;
        model &M = h->m;
        if (params)
        {
            for (int i = 0; i < P_COUNT; ++i) check_arg(std::isfinite(params[i]), "majorant parameters must be finite");
            gdx::hip_check(hipMemcpy(M.p.p, params, sizeof(float)*P_COUNT, hipMemcpyHostToDevice), "majorant parameters");
            gdx::hip_check(hipMemset(M.m1.p, 0, sizeof(float)*P_COUNT), "majorant parameters");
            gdx::hip_check(hipMemset(M.m2.p, 0, sizeof(float)*P_COUNT), "majorant parameters");
            M.steps = 0;
        }
        if (norm)
        {
            for (int k = 0; k < F; ++k)
            {
                check_arg(std::isfinite(norm[k]), "normalization mean must be finite");
                check_arg(std::isfinite(norm[F + k]) && norm[F + k] > 0.0f, "normalization scale must be positive");
                M.h_norm[k] = norm[k];
                M.h_norm[F + k] = 1.0f / norm[F + k];
            }
            gdx::hip_check(hipMemcpy(M.norm.p, M.h_norm.data(), sizeof(float)*2*F, hipMemcpyHostToDevice),
                           "majorant normalization");
            for (int k = 0; k < F; ++k) M.h_norm[F + k] = norm[F + k];
        }
        k_mlp_pack<<<1, 256>>>(M.p.p, M.norm.p, M.P.p);
        gdx::hip_check(hipGetLastError(), "majorant pack");
        gdx::hip_check(hipDeviceSynchronize(), "majorant pack");
        ++M.version;
    });
}

int gdx_col_majorant_infer(gdx_col_majorant h, const float *features, long long n, float *out)
{
    if (!h) return GDX_COL_ERR_ARG;
    return mguard(h, [&] {
        check_arg(n >= 0 && (n == 0 || (features && out)), "invalid inference arguments");
        gdx::mlp::launch_mlp_infer(h->m.P.p, features, n, out, nullptr);
        gdx::hip_check(hipDeviceSynchronize(), "majorant inference");
    });
}

int gdx_col_majorant_grad(gdx_col_majorant h, const float *features, const float *log_target, int n, float *grad,
                          double *loss)
{
    if (!h) return GDX_COL_ERR_ARG;
    return mguard(h, [&] {
        using namespace gdx::mlp;
        check_arg(n > 0 && features && log_target && grad, "invalid gradient arguments");
        model &M = h->m;
        int waves = grid_waves(n);
        int blocks = (waves + WAVES - 1) / WAVES;
        M.partial.alloc(size_t(blocks)*WAVES*P_STRIDE);
        k_mlp_grad<<<blocks, BLOCK>>>(M.P.p, features, log_target, nullptr, 0, n, M.partial.p);
        gdx::hip_check(hipGetLastError(), "majorant gradient launch");
        k_mlp_reduce<<<(P_LOSS + 256) / 256, 256>>>(M.partial.p, blocks*WAVES, M.grad.p);
        gdx::hip_check(hipGetLastError(), "majorant gradient launch");
        std::vector<float> h_grad(P_LOSS + 1);
        gdx::hip_check(hipMemcpy(h_grad.data(), M.grad.p, sizeof(float)*(P_LOSS + 1), hipMemcpyDeviceToHost),
                       "majorant gradient");
        std::memcpy(grad, h_grad.data(), sizeof(float)*P_COUNT);
        if (loss) *loss = double(h_grad[P_LOSS]) / n;
    });
}

int gdx_col_majorant_grad_noalloc(gdx_col_majorant h, const float *features, const float *log_target, int n,
                                  float *grad, double *loss)
{
    if (!h) return GDX_COL_ERR_ARG;
    return mguard(h, [&] {
        using namespace gdx::mlp;
        static_assert(16*TILES_PER_WAVE == 128 && GRAD_NOALLOC_ROWS % WAVES == 0, "scratch rows per wavefront");
        check_arg(n > 0 && features && log_target && grad, "invalid gradient arguments");
        check_arg(n <= GDX_COL_MAJORANT_GRAD_NOALLOC_MAX,
                  "gradient without allocation: n exceeds GDX_COL_MAJORANT_GRAD_NOALLOC_MAX samples");
        const int waves = grid_waves(n);
        const int blocks = (waves + WAVES - 1) / WAVES;
        const int rows = blocks*WAVES;
        if (rows > GRAD_NOALLOC_ROWS) throw gdx::status_error(GDX_COL_ERR_ARG, "scratch rows exceeded");
        float *partial = scratch_on_current_device();
        float *sum = partial + size_t(GRAD_NOALLOC_ROWS)*P_STRIDE;
        k_mlp_grad<<<blocks, BLOCK>>>(h->m.P.p, features, log_target, nullptr, 0, n, partial);
        gdx::hip_check(hipGetLastError(), "majorant gradient launch");
        k_mlp_reduce<<<(P_LOSS + 256) / 256, 256>>>(partial, rows, sum);
        gdx::hip_check(hipGetLastError(), "majorant gradient launch");
        float loss_sum = 0.0f;
        gdx::hip_check(hipMemcpy(grad, sum, sizeof(float)*P_COUNT, hipMemcpyDeviceToHost), "majorant gradient");
        gdx::hip_check(hipMemcpy(&loss_sum, sum + P_LOSS, sizeof(float), hipMemcpyDeviceToHost), "majorant gradient");
        if (loss) *loss = double(loss_sum) / n;
    });
}

int gdx_col_majorant_train(gdx_col_majorant h, const float *features, const float *log_target, long long n,
                           const gdx_col_train_config *cfg, float *epoch_loss)
{
    if (!h) return GDX_COL_ERR_ARG;
    return mguard(h, [&] {
        using namespace gdx::mlp;
        check_arg(cfg && features && log_target, "invalid training arguments");
        check_arg(cfg->batch > 0 && cfg->epochs >= 0 && n >= cfg->batch, "training needs at least one full minibatch");
        check_arg(n <= (1ll << 31) - 1, "training set too large");
        check_arg(std::isfinite(cfg->lr) && cfg->lr > 0.0, "learning rate must be positive");
        model &M = h->m;
        ++M.version;
        const int B = cfg->batch;
        const long long batches = n / B;               // the last partial minibatch is dropped
        int waves = grid_waves(B);
        int blocks = (waves + WAVES - 1) / WAVES;
        M.partial.alloc(size_t(blocks)*WAVES*P_STRIDE);
        M.perm.alloc(size_t(n));
        M.loss.alloc(size_t(std::max<long long>(1, batches)));
        std::vector<int> perm(static_cast<size_t>(n));
        std::vector<float> losses(static_cast<size_t>(std::max<long long>(1, batches)));
        uint64_t s = cfg->seed;
        const double b1 = cfg->beta1 > 0 ? cfg->beta1 : 0.9, b2 = cfg->beta2 > 0 ? cfg->beta2 : 0.999;
        const double eps = cfg->eps > 0 ? cfg->eps : 1e-8;
        for (int ep = 0; ep < cfg->epochs; ++ep)
        {
            // Fisher-Yates shuffle of the sample order, one per epoch
            for (long long i = 0; i < n; ++i) perm[size_t(i)] = int(i);
            for (long long i = n - 1; i > 0; --i)
            {
                long long k = (long long)(splitmix(s) % uint64_t(i + 1));
                std::swap(perm[size_t(i)], perm[size_t(k)]);
            }
            gdx::hip_check(hipMemcpy(M.perm.p, perm.data(), sizeof(int)*size_t(n), hipMemcpyHostToDevice),
                           "training order");
            for (long long b = 0; b < batches; ++b)
            {
                ++M.steps;
                adam_args A{float(cfg->lr), float(b1), float(b2), float(eps),
                            float(1.0 - std::pow(b1, double(M.steps))), float(1.0 - std::pow(b2, double(M.steps)))};
                k_mlp_grad<<<blocks, BLOCK>>>(M.P.p, features, log_target, M.perm.p, b*B, B, M.partial.p);
                k_mlp_adam<<<1, 1024>>>(M.p.p, M.m1.p, M.m2.p, M.partial.p, blocks*WAVES, A, M.norm.p, M.P.p,
                                        M.loss.p + b);
            }
            gdx::hip_check(hipGetLastError(), "training launch");
            gdx::hip_check(hipMemcpy(losses.data(), M.loss.p, sizeof(float)*size_t(batches), hipMemcpyDeviceToHost),
                           "training loss");
            if (epoch_loss)
            {
                double sum = 0.0;
                for (long long b = 0; b < batches; ++b) sum += losses[size_t(b)];
                epoch_loss[ep] = float(sum / double(batches*B));
            }
        }
        gdx::hip_check(hipDeviceSynchronize(), "training");
    });
}

}  // extern "C"

namespace gdx {

const void *majorant_pack(gdx_col_majorant m) { return m ? static_cast<const void *>(m->m.P.p) : nullptr; }
unsigned long long majorant_version(gdx_col_majorant m) { return m ? m->m.version : 0ull; }

}  // namespace gdx
