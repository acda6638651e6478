// encode.hip -- the CAQ encoder on gfx950 (IVF::construct's quantisation,
// saqlib/index/ivf.hpp:161-183 -> SAQuantizer::quantize_cluster, saq_quantizer.hpp:34-60).
//
//   rotate_kernel  residual o = x_s * P_s - c_rot (quantizer.hpp:45-48) for a chunk of vectors,
//                  each row accumulated in the order of Eigen's GEMV (FMA chains over blocks of
//                  16 input dims, summed block by block), written transposed [D_s][chunk] so
//                  the encoder's per-vector sweeps are coalesced.
//   caq_kernel     one thread per vector: CAQEncoder::encode (max |o|, floor codes with the
//                  release build's hoisted reciprocal), the FP64 code-adjustment loop in
//                  source order, encode_and_fac's factors and rescale_vmx_to1, then the
//                  ClusterPacker output: the MSB bit of each code into the vector's lane of
//                  its block record, the low bits into the CodeHelper<b-1> compacted long
//                  code, ExFactor and the block factors (caq_encoder.hpp:57-247,
//                  cluster_packer.hpp:59-139, code_helper.hpp:58-568).
#include <cfloat>

#include "common.hpp"

#pragma clang fp contract(off)

namespace saqgpu {

struct EncArgs {
    IndexView ix;
    const float *data;     // rows of num_dim floats
    const uint32_t *src;   // per slot: source row
    const int64_t *gvec;   // per slot: destination vector index
    const int64_t *gblk;   // per slot: destination block
    const int32_t *lane;   // per slot: lane in that block
    const uint32_t *clu;   // per slot: cluster
    const float *cent;     // K x num_dim unrotated centroids
    const float *rot;      // P_s or nullptr
    int32_t seg;
    int64_t nslots;
    int64_t ld;            // leading dimension (slots) of the transposed scratch
    int32_t rd_lmt;
    float eps;
    int32_t ori_qb;        // caq_ori_qB: 0, or the bits the codes are found at before the shift down
    float *O;              // [D_s][ld]
    uint16_t *code;        // [D_s][ld]
    int *fail;
    // SaqSingleDataWrapper records instead of the index's lists (src, clu, gvec, gblk, lane and
    // cent unused; slot t is row t and record t; no centroid)
    uint8_t *srec;
    int64_t sstride;
    int32_t s_code_off, s_long_off, s_exf_off;
    int64_t cent_stride = -1;  // floats between lists' centroids in cent (0: one shared centroid)
};

constexpr int kRotRows = 8;

// grid (ceil(nslots / 8), ceil(D_s / 256))
__global__ __launch_bounds__(256) void rotate_kernel(EncArgs a) {
    __shared__ float xs[kRotRows][256];
    const IndexView &ix = a.ix;
    const SegDesc sd = ix.segs[a.seg];
    const int D = sd.dim;
    const int j = blockIdx.y * 256 + threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * kRotRows;
    const int bc = D < 128 ? D : (D * 4 < 32000 ? 16 : 4);
    float acc[kRotRows], c[kRotRows];
#pragma unroll
    for (int t = 0; t < kRotRows; ++t) acc[t] = c[t] = 0.0f;
    if (a.rot) {
        for (int i0 = 0; i0 < D; i0 += 256) {
            __syncthreads();
            for (int t = 0; t < kRotRows; ++t) {
                const int64_t r = r0 + t;
                const int d = sd.doff + i0 + threadIdx.x;
                float v = 0.0f;
                if (r < a.nslots && i0 + (int)threadIdx.x < D && d < ix.num_dim)
                    v = a.data[(a.src ? (int64_t)a.src[r] : r) * ix.num_dim + d];
                xs[t][threadIdx.x] = v;
            }
            __syncthreads();
            if (j < D) {
                const int iend = min(256, D - i0);
                for (int ii = 0; ii < iend; ++ii) {
                    const int i = i0 + ii;
                    const float p = a.rot[(int64_t)i * D + j];
#pragma unroll
                    for (int t = 0; t < kRotRows; ++t) c[t] = __fmaf_rn(p, xs[t][ii], c[t]);
                    if ((i + 1) % bc == 0 || i + 1 == D) {
#pragma unroll
                        for (int t = 0; t < kRotRows; ++t) {
                            acc[t] = acc[t] + c[t];
                            c[t] = 0.0f;
                        }
                    }
                }
            }
        }
    }
    if (j >= D) return;
    for (int t = 0; t < kRotRows; ++t) {
        const int64_t r = r0 + t;
        if (r >= a.nslots) break;
        const int d = sd.doff + j;
        float x = acc[t];
        if (!a.rot) x = d < ix.num_dim ? a.data[(a.src ? (int64_t)a.src[r] : r) * ix.num_dim + d] : 0.0f;
        a.O[(int64_t)j * a.ld + r] = a.srec ? x : x - ix.crot[(int64_t)a.clu[r] * ix.D_pad + d];
    }
}

// The same rotation with lanes over vectors: a workgroup takes 256 R vectors (4 waves; lane l
// holds vectors l, l + 256, ...) and walks the output columns in blocks of JB; for each block it
// steps over the input dims 16 at a time.  A step stages in LDS the 16 x JB tile of P, read back
// as broadcasts so that each P value feeds R vectors' FMAs, and the 16 dims of the workgroup's
// vectors, loaded 64-byte row pieces at a time (lanes owning whole rows would touch 64 rows per
// load).  Every column is the FMA chain of Eigen's GEMV over blocks of BC input dims (BC = 0: one
// block of all D dims, for D < 128), the blocks' chains summed in order -- the sums rotate_kernel
// forms, bit for bit.  The 16 dims of a step go G at a time through a runtime loop, which bounds
// the P values the scheduler holds in registers to G JB.  Requires D a multiple of 64 (segment
// dims are).  grid: ceil(nslots / (256 R)).
// (JB 64, R 1, G 4: the fastest of JB 16..64, R 1..2, G 2..4 on 10M x 1024 vectors)
constexpr int kRotJB = 64, kRotR = 1, kRotG = 4;
constexpr int kRotXS = 20;  // floats per staged row: 16 dims + 4 (conflict-free float4 reads)

template <int BC>
__global__ __launch_bounds__(256) void rotate_rows_kernel(EncArgs a) {
    constexpr int JB = kRotJB, R = kRotR, G = kRotG;
    constexpr int NR = 256 * R;  // vectors of the workgroup
    __shared__ __align__(16) float tile[16][JB];
    __shared__ __align__(16) float xs[NR][kRotXS];
    const IndexView &ix = a.ix;
    const SegDesc sd = ix.segs[a.seg];
    const int D = sd.dim;
    const int nd = min(D, ix.num_dim - sd.doff);  // dims of the segment inside the data rows
    const bool vec4 = (ix.num_dim % 4 == 0) && (sd.doff % 4 == 0);
    const int64_t v0 = (int64_t)blockIdx.x * NR;
    // staging of the vectors' dims: thread t moves the 4 dims 4 (t % 4) .. + 3 of rows t / 4 + 64 q
    const float *srow[4 * R];
    bool slive[4 * R];
#pragma unroll
    for (int q = 0; q < 4 * R; ++q) {
        const int64_t r = v0 + threadIdx.x / 4 + 64 * q;
        slive[q] = r < a.nslots;
        const int64_t row = slive[q] ? (a.src ? (int64_t)a.src[r] : r) : 0;
        srow[q] = a.data + row * ix.num_dim + sd.doff + 4 * (threadIdx.x % 4);
    }
    const int sc = 4 * (threadIdx.x % 4);
    const float *P = a.rot;
    constexpr int kStage = (16 * JB + 255) / 256;
    const int nsteps = D / 16;
    for (int j0 = 0; j0 < D; j0 += JB) {
        float acc[R][JB], c[R][JB];
#pragma unroll
        for (int k = 0; k < R; ++k)
#pragma unroll
            for (int j = 0; j < JB; ++j) acc[k][j] = c[k][j] = 0.0f;
        for (int k = 0; k < nsteps; ++k) {
            const int i0 = 16 * k;
            // this step's P tile and vector dims, loaded before the barrier that ends the previous
            // step's reads, stored after it
            float pn[kStage];
#pragma unroll
            for (int q = 0; q < kStage; ++q) {
                const int e = threadIdx.x + 256 * q;
                if (e < 16 * JB) pn[q] = P[(int64_t)(i0 + e / JB) * D + j0 + e % JB];
            }
            float4 xn[4 * R];
#pragma unroll
            for (int q = 0; q < 4 * R; ++q) {
                const int d = i0 + sc;
                if (slive[q] && vec4 && d + 4 <= nd) {
                    xn[q] = *reinterpret_cast<const float4 *>(srow[q] + i0);
                } else {
                    xn[q].x = (slive[q] && d < nd) ? srow[q][i0] : 0.0f;
                    xn[q].y = (slive[q] && d + 1 < nd) ? srow[q][i0 + 1] : 0.0f;
                    xn[q].z = (slive[q] && d + 2 < nd) ? srow[q][i0 + 2] : 0.0f;
                    xn[q].w = (slive[q] && d + 3 < nd) ? srow[q][i0 + 3] : 0.0f;
                }
            }
            __syncthreads();
#pragma unroll
            for (int q = 0; q < kStage; ++q) {
                const int e = threadIdx.x + 256 * q;
                if (e < 16 * JB) tile[e / JB][e % JB] = pn[q];
            }
#pragma unroll
            for (int q = 0; q < 4 * R; ++q)
                *reinterpret_cast<float4 *>(&xs[threadIdx.x / 4 + 64 * q][sc]) = xn[q];
            __syncthreads();
#pragma unroll 1
            for (int g = 0; g < 16 / G; ++g) {
                float xv[R][G];
#pragma unroll
                for (int q = 0; q < R; ++q) {
                    const float *xr = &xs[threadIdx.x + 256 * q][G * g];
                    if (G == 4) {
                        const float4 v = *reinterpret_cast<const float4 *>(xr);
                        xv[q][0] = v.x;
                        xv[q][1 % G] = v.y;
                        xv[q][2 % G] = v.z;
                        xv[q][3 % G] = v.w;
                    } else {
#pragma unroll
                        for (int u = 0; u < G; ++u) xv[q][u] = xr[u];
                    }
                }
#pragma unroll
                for (int ii = 0; ii < G; ++ii) {
#pragma unroll
                    for (int j4 = 0; j4 < JB; j4 += 4) {
                        const float4 p = *reinterpret_cast<const float4 *>(&tile[G * g + ii][j4]);
#pragma unroll
                        for (int q = 0; q < R; ++q) {
                            c[q][j4] = __fmaf_rn(p.x, xv[q][ii], c[q][j4]);
                            c[q][j4 + 1] = __fmaf_rn(p.y, xv[q][ii], c[q][j4 + 1]);
                            c[q][j4 + 2] = __fmaf_rn(p.z, xv[q][ii], c[q][j4 + 2]);
                            c[q][j4 + 3] = __fmaf_rn(p.w, xv[q][ii], c[q][j4 + 3]);
                        }
                    }
                    // Eigen's block ends: every 4 dims (BC 4) or 16 (BC 16)
                    const int idx = G * g + ii;
                    if ((BC == 4 && idx % 4 == 3) || (BC == 16 && idx == 15)) {
#pragma unroll
                        for (int q = 0; q < R; ++q)
#pragma unroll
                            for (int j = 0; j < JB; ++j) {
                                acc[q][j] = acc[q][j] + c[q][j];
                                c[q][j] = 0.0f;
                            }
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < R; ++q) {
            if (BC == 0)
#pragma unroll
                for (int j = 0; j < JB; ++j) acc[q][j] = acc[q][j] + c[q][j];
            const int64_t r = v0 + threadIdx.x + 256 * q;
            if (r >= a.nslots) continue;
#pragma unroll
            for (int j = 0; j < JB; ++j) {
                const int d = sd.doff + j0 + j;
                a.O[(int64_t)(j0 + j) * a.ld + r] =
                    a.srec ? acc[q][j] : acc[q][j] - ix.crot[(int64_t)a.clu[r] * ix.D_pad + d];
            }
        }
    }
}

// Eigen redux (Redux.h:227-272) of n elements (n a multiple of 64) by one thread.
template <class F>
__device__ __forceinline__ float redux_thread64(int n, F elem) {
    float p0[16], p1[16];
#pragma unroll
    for (int l = 0; l < 16; ++l) {
        p0[l] = elem(l);
        p1[l] = elem(16 + l);
    }
    for (int idx = 32; idx < n; idx += 32) {
#pragma unroll
        for (int l = 0; l < 16; ++l) {
            p0[l] = p0[l] + elem(idx + l);
            p1[l] = p1[l] + elem(idx + 16 + l);
        }
    }
    float b[8], c[4];
#pragma unroll
    for (int l = 0; l < 16; ++l) p0[l] = p0[l] + p1[l];
#pragma unroll
    for (int i = 0; i < 8; ++i) b[i] = p0[i] + p0[i + 8];
#pragma unroll
    for (int i = 0; i < 4; ++i) c[i] = b[i] + b[i + 4];
    return (c[0] + c[2]) + (c[1] + c[3]);
}

// One byte of the CodeHelper<k> compacted layout (code_helper.hpp:95-568) for k = 1..8.
// low(d) yields the low k bits of the code of dim d.
template <class F>
__device__ __forceinline__ void compact_store(int k, int D, uint8_t *out, F low) {
    switch (k) {
    case 1:
        for (int i = 0; i < D; i += 8) {
            uint32_t v = 0;
            for (int t = 0; t < 8; ++t) v |= (low(i + t) & 1u) << t;
            out[i / 8] = (uint8_t)v;
        }
        break;
    case 2:
    case 3: {
        const int stride = k == 2 ? 16 : 24;
        for (int c = 0; c < D / 64; ++c) {
            uint8_t *o = out + stride * c;
            for (int b = 0; b < 16; ++b) {
                uint32_t v = 0;
                for (int p = 0; p < 4; ++p) v |= (low(64 * c + 16 * p + b) & 3u) << (2 * p);
                o[b] = (uint8_t)v;
            }
            if (k == 3)
                for (int kk = 0; kk < 8; ++kk) {
                    uint32_t v = 0;
                    for (int t = 0; t < 8; ++t) v |= ((low(64 * c + 8 * t + kk) >> 2) & 1u) << t;
                    o[16 + kk] = (uint8_t)v;
                }
        }
        break;
    }
    case 4:
        for (int c = 0; c < D / 32; ++c)
            for (int b = 0; b < 16; ++b)
                out[16 * c + b] = (uint8_t)((low(32 * c + b) & 15u) | ((low(32 * c + 16 + b) & 15u) << 4));
        break;
    case 6:
    case 7: {
        const int stride = k == 6 ? 48 : 56;
        for (int c = 0; c < D / 64; ++c) {
            uint8_t *o = out + stride * c;
            for (int b = 0; b < 16; ++b) {
                const uint32_t v0 = low(64 * c + b), v1 = low(64 * c + 16 + b);
                const uint32_t v2 = low(64 * c + 32 + b), v3 = low(64 * c + 48 + b);
                o[b] = (uint8_t)((v0 & 0x3F) | ((v2 << 2) & 0xC0));
                o[16 + b] = (uint8_t)((v1 & 0x3F) | ((v3 << 2) & 0xC0));
                o[32 + b] = (uint8_t)((v2 & 0x0F) | ((v3 & 0x0F) << 4));
            }
            if (k == 7)
                for (int kk = 0; kk < 8; ++kk) {
                    uint32_t v = 0;
                    for (int t = 0; t < 8; ++t) v |= ((low(64 * c + 8 * t + kk) >> 6) & 1u) << t;
                    o[48 + kk] = (uint8_t)v;
                }
        }
        break;
    }
    case 8:
        for (int i = 0; i < D; ++i) out[i] = (uint8_t)low(i);
        break;
    default:
        break;
    }
}

template <class F>
__device__ __forceinline__ void compact_code(int k, int D, uint8_t *out, F low) {
    if (k == 0) return;
    if (k == 5) {  // CodeHelper<4> of code >> 1, then CodeHelper<1> of code & 1
        compact_store(4, D, out, [&](int d) { return low(d) >> 1; });
        compact_store(1, D, out + D * 4 / 8, [&](int d) { return low(d) & 1u; });
        return;
    }
    if (k > 8) {  // low bytes, then CodeHelper<k-8> of the high bits
        for (int i = 0; i < D; ++i) out[i] = (uint8_t)(low(i) & 0xFF);
        const int kh = k - 8;
        if (kh == 5) {
            compact_store(4, D, out + D, [&](int d) { return (low(d) >> 8) >> 1; });
            compact_store(1, D, out + D + D * 4 / 8, [&](int d) { return (low(d) >> 8) & 1u; });
        } else {
            compact_store(kh, D, out + D, [&](int d) { return low(d) >> 8; });
        }
        return;
    }
    compact_store(k, D, out, low);
}

__global__ __launch_bounds__(256) void caq_kernel(EncArgs a) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= a.nslots) return;
    const IndexView &ix = a.ix;
    const SegDesc sd = ix.segs[a.seg];
    const int D = sd.dim, bits = sd.bits;
    const int64_t ld = a.ld;
    const float *o = a.O + t;
    uint16_t *code = a.code + t;
    auto O = [&](int j) { return o[(int64_t)j * ld]; };

    double v_mx = 0, v_mi = 0, delta = 0, ip = 0, oa_l2sqr = 0;
    const int cm = (1 << (a.ori_qb ? a.ori_qb : bits)) - 1;
    // one pass over the residual: its max and min, and Eigen's redux (Redux.h:227-272) of o and
    // of o * o (vec_sum, and |o|^2 for encode_and_fac) in redux_thread64's element order
    float mx, mn, o_sum, o_sq;
    {
        float p0[16], p1[16], r0[16], r1[16];
#pragma unroll
        for (int l = 0; l < 16; ++l) {
            const float x0 = O(l), x1 = O(16 + l);
            p0[l] = x0;
            p1[l] = x1;
            r0[l] = x0 * x0;
            r1[l] = x1 * x1;
        }
        mx = mn = p0[0];
#pragma unroll
        for (int l = 0; l < 16; ++l) {
            mx = fmaxf(fmaxf(mx, p0[l]), p1[l]);
            mn = fminf(fminf(mn, p0[l]), p1[l]);
        }
        for (int idx = 32; idx < D; idx += 32) {
#pragma unroll
            for (int l = 0; l < 16; ++l) {
                const float x0 = O(idx + l), x1 = O(idx + 16 + l);
                mx = fmaxf(fmaxf(mx, x0), x1);
                mn = fminf(fminf(mn, x0), x1);
                p0[l] = p0[l] + x0;
                p1[l] = p1[l] + x1;
                r0[l] = r0[l] + x0 * x0;
                r1[l] = r1[l] + x1 * x1;
            }
        }
        float b[8], c[4], br[8], cr[4];
#pragma unroll
        for (int l = 0; l < 16; ++l) {
            p0[l] = p0[l] + p1[l];
            r0[l] = r0[l] + r1[l];
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            b[i] = p0[i] + p0[i + 8];
            br[i] = r0[i] + r0[i + 8];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            c[i] = b[i] + b[i + 4];
            cr[i] = br[i] + br[i + 4];
        }
        o_sum = (c[0] + c[2]) + (c[1] + c[3]);
        o_sq = (cr[0] + cr[2]) + (cr[1] + cr[3]);
    }
    if (bits > 0) {
        v_mx = (double)fmaxf(mx, -mn);
        v_mi = -v_mx;
        delta = (v_mx - v_mi) / (cm + 1);
        const double vec_sum = (double)o_sum;
        const float fvmi = (float)v_mi, finv = 1.0f / (float)delta, fcm = (float)cm;
        double ip_o_code = 0;
        unsigned long long code_l2sqr = 0;
        uint32_t code_sum = 0;
        for (int j = 0; j < D; ++j) {
            const float oj = O(j);
            float cf = 0.0f;
            if (delta != 0) {
                const float fl = floorf((oj - fvmi) * finv);
                cf = (fcm < fl) ? fcm : fl;  // std::min(floor, code_max)
            }
            ip_o_code += (double)(cf * oj);
            const int c = (int)cf;
            code[(int64_t)j * ld] = (uint16_t)c;
            code_l2sqr += (unsigned long long)(c * c);
            code_sum += (uint32_t)c;
        }
        ip = ip_o_code * delta + (v_mi + 0.5 * delta) * vec_sum;
        oa_l2sqr = delta * delta * (double)code_l2sqr + (delta * delta + 2 * delta * v_mi) * (double)code_sum;
        oa_l2sqr += (0.25 * delta * delta + delta * v_mi + v_mi * v_mi) * (double)D;
        if (a.rd_lmt && oa_l2sqr != 0) {  // code_adjustment, caq_encoder.hpp:57-142
            const double re_eps = (double)a.eps * oa_l2sqr;
            for (int adj_cnt = 1, round = 1; adj_cnt; round++) {
                if (!(ip >= 0)) {
                    atomicAdd(a.fail, 1);  // the reference would abort on CHECK_GE(ip_resi_oa, 0)
                    break;
                }
                adj_cnt = 0;
                // the exact recompute between rounds (caq_encoder.hpp:126-140) summed as each
                // dim's code becomes final for the round: the same terms in the same order
                double chk_oa = 0, chk_ip = 0;
                // 8 dims' residuals and codes are loaded together ahead of the sequential updates
                // (D is a multiple of 64)
                for (int j8 = 0; j8 < D; j8 += 8) {
                  float ov[8];
                  int cv[8];
#pragma unroll
                  for (int u = 0; u < 8; ++u) {
                      ov[u] = O(j8 + u);
                      cv[u] = code[(int64_t)(j8 + u) * ld];
                  }
#pragma unroll
                  for (int u = 0; u < 8; ++u) {
                    const int j = j8 + u;
                    int cur = 0;
                    const float oj = ov[u];
                    const int c0 = cv[u];
                    double oa = (c0 + 0.5) * delta + v_mi;
                    int c = c0;
                    const double tmp = oa_l2sqr - oa * oa;
                    const double ip_delta = delta * oj;
                    while (c < cm) {
                        const double nq = oa + delta, nl = tmp + nq * nq, ni = ip + ip_delta;
                        if ((ip * ip + re_eps) * nl >= ni * ni * oa_l2sqr) break;
                        c++;
                        ip = ni;
                        oa = nq;
                        oa_l2sqr = nl;
                        cur++;
                    }
                    while (c > 0) {
                        const double nq = oa - delta, nl = tmp + nq * nq, ni = ip - ip_delta;
                        if ((ip * ip + re_eps) * nl >= ni * ni * oa_l2sqr) break;
                        c--;
                        ip = ni;
                        oa = nq;
                        oa_l2sqr = nl;
                        cur++;
                    }
                    if (c0 != c) {
                        code[(int64_t)j * ld] = (uint16_t)c;
                        adj_cnt += cur;
                    }
                    const double q = (c + 0.5) * delta + v_mi;
                    chk_ip += q * (double)oj;
                    chk_oa += q * q;
                  }
                }
                if (a.rd_lmt && round >= a.rd_lmt) break;
                oa_l2sqr = chk_oa;
                ip = chk_ip;
            }
        }
        if (a.ori_qb) {  // CAQEncoder::downUpSample (caq_encoder.hpp:144-157)
            const int shift = a.ori_qb - bits;
            delta *= 1 << shift;
            ip = 0;
            oa_l2sqr = 0;
            for (int j = 0; j < D; ++j) {
                const int c = code[(int64_t)j * ld] >> shift;
                code[(int64_t)j * ld] = (uint16_t)c;
                const double q = (c + 0.5) * delta + v_mi;
                ip += q * (double)O(j);
                oa_l2sqr += q * q;
            }
        }
    }
    // encode_and_fac (caq_encoder.hpp:222-247)
    const double o_l2sqr = (double)o_sq;
    const double o_l2norm = sqrt(o_l2sqr);
    double fac_rescale = ip != 0 ? o_l2sqr / ip : 0;
    const double fac_error = o_l2sqr * (double)1.9f * sqrt((((o_l2sqr * oa_l2sqr) / (ip * ip)) - 1) / (double)(D - 1));
    if (v_mx != 0) {  // rescale_vmx_to1
        const double sr = 1.0 / v_mx;
        delta *= sr;
        ip *= sr;
        oa_l2sqr *= sr * sr;
        fac_rescale *= v_mx;
        v_mi *= sr;
        v_mx *= sr;
    }
    if (a.srec) {  // QuantizerSingle::store_quantization_result (quantizer.hpp:134-207)
        const uint32_t short_bit = bits ? 1u << (bits - 1) : 0u;
        uint8_t *rec = a.srec + t * a.sstride;
        float *fac = reinterpret_cast<float *>(rec) + 2 * a.seg;
        fac[0] = (float)o_l2norm;
        if (bits == 0) return;
        fac[1] = 0.0f;  // ip_cent_oa: no centroid
        float *ef = reinterpret_cast<float *>(rec + a.s_exf_off) + 2 * a.seg;
        ef[0] = (float)fac_rescale;
        ef[1] = (float)fac_error;
        // MSB-first bytes, the bytes of every 8 reversed (:166-188): byte j of a group of 8 lands
        // in byte 7 - j of a little-endian word
        uint64_t *sc = reinterpret_cast<uint64_t *>(rec + a.s_code_off);
        for (int g = 0; g < D / 64; ++g) {
            uint64_t w = 0;
            for (int b = 0; b < 8; ++b) {
                uint32_t byte = 0;
                for (int tt = 0; tt < 8; ++tt)
                    byte |= ((code[(int64_t)(64 * g + 8 * b + tt) * ld] & short_bit) ? 0x80u : 0u) >> tt;
                w |= (uint64_t)byte << (8 * (7 - b));
            }
            sc[g] = w;
        }
        if (bits > 1)
            compact_code(bits - 1, D, rec + a.s_long_off,
                         [&](int d) { return (uint32_t)code[(int64_t)d * ld] & (short_bit - 1u); });
        return;
    }
    const int64_t gv = a.gvec[t], gb = a.gblk[t];
    const int ln = a.lane[t];
    uint8_t *rec = const_cast<uint8_t *>(ix.blocks) + gb * ix.block_stride;
    reinterpret_cast<float *>(rec)[a.seg * kBlock + ln] = (float)o_l2norm;
    if (bits == 0) return;
    // ip_cent_oa = centroid . get_oa() with the unrotated segment centroid (caq_encoder.hpp:42-45, 241)
    const float fd = (float)delta, fm = (float)v_mi;
    const float *cen = a.cent + (int64_t)a.clu[t] * (a.cent_stride < 0 ? ix.num_dim : a.cent_stride);
    const float ipc = redux_thread64(D, [&](int j) {
        const int d = sd.doff + j;
        const float cj = d < ix.num_dim ? cen[d] : 0.0f;
        return cj * ((float)code[(int64_t)j * ld] * fd + fm);
    });
    const_cast<float *>(ix.ipcent)[(gb * ix.nseg + a.seg) * kBlock + ln] = ipc;
    float *ef = const_cast<float *>(ix.exfac) + (gv * ix.nseg + a.seg) * 2;
    ef[0] = (float)fac_rescale;
    ef[1] = (float)fac_error;
    // short code: MSB of each code, MSB-first bytes (cluster_packer.hpp:124-139)
    const uint32_t short_bit = 1u << (bits - 1);
    uint8_t *codes = rec + (int64_t)ix.nseg * kBlock * 4;
    for (int g = 0; g < sd.ngroups; ++g) {
        uint32_t w[2] = {0, 0};
        for (int b = 0; b < 8; ++b) {
            uint32_t byte = 0;
            for (int tt = 0; tt < 8; ++tt)
                byte |= ((code[(int64_t)(64 * g + 8 * b + tt) * ld] & short_bit) ? 0x80u : 0u) >> tt;
            w[b / 4] |= byte << (8 * (b % 4));
        }
        uint2 *dst = reinterpret_cast<uint2 *>(codes + (int64_t)(sd.group_off + g) * 256 + ln * 8);
        *dst = make_uint2(w[0], w[1]);
    }
    // long code: low bits in the CodeHelper<bits-1> layout
    uint8_t *lc = const_cast<uint8_t *>(ix.long_codes) + gv * ix.long_stride + sd.long_off;
    compact_code(bits - 1, D, lc, [&](int d) { return (uint32_t)code[(int64_t)d * ld] & (short_bit - 1u); });
}

// Residuals of a chunk: the lane-per-vector kernel where P is given (Eigen's blocking of the
// input dims: all of them below 128, else 16, or 4 from 8000 on), the column kernel otherwise.
static void launch_rotate(const EncArgs &a, int seg_dim, hipStream_t st) {
    if (a.rot && seg_dim % 64 == 0) {
        const dim3 g((unsigned)((a.nslots + 256 * kRotR - 1) / (256 * kRotR)));
        if (seg_dim < 128)
            hipLaunchKernelGGL(rotate_rows_kernel<0>, g, dim3(256), 0, st, a);
        else if (seg_dim * 4 < 32000)
            hipLaunchKernelGGL(rotate_rows_kernel<16>, g, dim3(256), 0, st, a);
        else
            hipLaunchKernelGGL(rotate_rows_kernel<4>, g, dim3(256), 0, st, a);
        return;
    }
    const dim3 g1((unsigned)((a.nslots + kRotRows - 1) / kRotRows), (unsigned)((seg_dim + 255) / 256));
    hipLaunchKernelGGL(rotate_kernel, g1, dim3(256), 0, st, a);
}

hipError_t launch_encode_segment(const IndexView &ix, const float *data, const uint32_t *src, const int64_t *gvec,
                                 const int64_t *gblk, const int32_t *lane, const uint32_t *clu, const float *cent,
                                 const float *rot, int seg, int seg_dim, int64_t nslots, int64_t ld, int rd_lmt,
                                 float eps, int ori_qb, bool shared_cent, float *O, uint16_t *code, int *fail,
                                 hipStream_t st) {
    EncArgs a{ix, data, src, gvec, gblk, lane, clu, cent, rot, seg, nslots, ld, rd_lmt, eps, ori_qb, O, code, fail};
    a.cent_stride = shared_cent ? 0 : -1;
    launch_rotate(a, seg_dim, st);
    hipLaunchKernelGGL(caq_kernel, dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_single_quantize_segment(const IndexView &ix, const float *data, const float *rot, int seg,
                                          int seg_dim, int64_t nslots, int64_t ld, int rd_lmt, float eps, int ori_qb,
                                          float *O, uint16_t *code, int *fail, uint8_t *srec, int64_t sstride,
                                          int32_t code_off, int32_t long_off, int32_t exf_off, hipStream_t st) {
    EncArgs a{ix, data, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, rot, seg, nslots, ld, rd_lmt, eps,
              ori_qb, O, code, fail, srec, sstride, code_off, long_off, exf_off};
    launch_rotate(a, seg_dim, st);
    hipLaunchKernelGGL(caq_kernel, dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_caq_only(const IndexView &ix, const int64_t *gvec, const int64_t *gblk, const int32_t *lane,
                           const uint32_t *clu, const float *cent, int seg, int64_t nslots, int64_t ld, int rd_lmt,
                           float eps, int ori_qb, float *O, uint16_t *code, int *fail, hipStream_t st) {
    EncArgs a{ix, nullptr, nullptr, gvec, gblk, lane, clu, cent, nullptr, seg, nslots, ld, rd_lmt, eps, ori_qb, O, code,
              fail};
    hipLaunchKernelGGL(caq_kernel, dim3((unsigned)((nslots + 255) / 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace saqgpu
