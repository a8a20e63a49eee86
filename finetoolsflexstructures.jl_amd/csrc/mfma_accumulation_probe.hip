// FP64 MFMA against the VALU for the block accumulation of the Q4RS element
// stiffness, K_e = sum_q C_q^T C_q with C_q 8 x 24 per Gauss point, 4 Gauss
// points, 10 upper 6x6 node blocks (FEMMShellQ4RSModule.jl:930-934 forms the
// same products as dense B^T D B), on one MI355X.  Prints:
//   1. the lane layouts of v_mfma_f64_16x16x4f64 and v_mfma_f64_4x4x4f64
//      (4 blocks) deduced from one-hot products, and a check that the
//      4x4x4 layout written into the accumulation kernel below is that one;
//   2. issue cost per wave instruction of both MFMAs and of v_fma_f64;
//   3. the accumulation over 262,144 elements: VALU (one lane per (element,
//      upper block), 8 rows of C from LDS per Gauss point, as the element
//      kernel does) and MFMA 4x4x4 (one wave per 6 elements, the 21 upper
//      4x4 tiles of the 24 x 24 result), each minus the fill of C that both
//      share, and the two results compared.  The C values carry 16
//      significant bits, so every product and sum is exact and the two
//      orders of summation must agree bit for bit: the check covers the
//      MFMA operand and result indexing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

#define CK(x)                                                                              \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);    \
            std::exit(1);                                                                  \
        }                                                                                  \
    } while (0)

typedef double d4 __attribute__((ext_vector_type(4)));

// ---- layout deduction: A one-hot at lane blockIdx.x, B[lane] = lane + 1 ----
__global__ void onehot16(double *out) {
    const int l = threadIdx.x;
    const double a = (l == (int)blockIdx.x) ? 1.0 : 0.0, b = l + 1.0;
    d4 c = {0.0, 0.0, 0.0, 0.0};
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    for (int v = 0; v < 4; ++v) out[(blockIdx.x * 64 + l) * 4 + v] = c[v];
}
__global__ void onehot4(double *out) {
    const int l = threadIdx.x;
    const double a = (l == (int)blockIdx.x) ? 1.0 : 0.0, b = l + 1.0;
    double c = 0.0;
    c = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
    out[blockIdx.x * 64 + l] = c;
}

// ---- issue cost: independent chains per wave ----
template <int KIND>
__global__ void issue(int iters, double *out) {
    const int l = threadIdx.x;
    double a = 1.0 + 1e-9 * l, b = 1.0 - 1e-9 * l;
    if (KIND == 0) {
        d4 c0 = {0, 0, 0, 0}, c1 = c0, c2 = c0, c3 = c0;
        for (int i = 0; i < iters; ++i) {
            c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c0, 0, 0, 0);
            c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c1, 0, 0, 0);
            c2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c2, 0, 0, 0);
            c3 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c3, 0, 0, 0);
        }
        out[blockIdx.x * blockDim.x + l] = c0[0] + c1[1] + c2[2] + c3[3];
    } else if (KIND == 1) {
        double c0 = 0, c1 = 0, c2 = 0, c3 = 0;
        for (int i = 0; i < iters; ++i) {
            c0 = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c0, 0, 0, 0);
            c1 = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c1, 0, 0, 0);
            c2 = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c2, 0, 0, 0);
            c3 = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c3, 0, 0, 0);
        }
        out[blockIdx.x * blockDim.x + l] = c0 + c1 + c2 + c3;
    } else {
        double c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0, c5 = 0, c6 = 0, c7 = 0;
        for (int i = 0; i < iters; ++i) {
            c0 = fma(a, b, c0); c1 = fma(a, b, c1); c2 = fma(a, b, c2); c3 = fma(a, b, c3);
            c4 = fma(b, a, c4); c5 = fma(b, a, c5); c6 = fma(b, a, c6); c7 = fma(b, a, c7);
            a = a * 1.0000000001;
        }
        out[blockIdx.x * blockDim.x + l] = c0 + c1 + c2 + c3 + c4 + c5 + c6 + c7;
    }
}

// ---- the accumulation ----
// C of element e, Gauss point q, row s, column col: 16 significant bits
__device__ __forceinline__ double cval(int e, int q, int s, int col) {
    unsigned h = (unsigned)e * 2654435761u ^ (unsigned)(q * 8 + s) * 40503u ^ (unsigned)col * 2246822519u;
    h ^= h >> 13;
    h *= 0x5bd1e995u;
    h ^= h >> 15;
    return (double)(h & 0xffff) / 65536.0 - 0.5;
}
constexpr int NE = 12, WG = 128, CSTR = 8 * 24 + 2;

__device__ __forceinline__ void fill_c(double *sC, int e0, int q) {
    for (int i = threadIdx.x; i < NE * 8 * 24; i += WG) {
        const int le = i / 192, rem = i % 192;
        sC[le * CSTR + rem] = cval(e0 + le, q, rem / 24, rem % 24);
    }
}

// the fill alone (the cost both variants share)
__global__ __launch_bounds__(WG) void acc_fill(int nelem, double *out) {
    __shared__ double sC[NE * CSTR];
    const int tid = threadIdx.x, e0 = blockIdx.x * NE;
    double s = 0.0;
    for (int q = 0; q < 4; ++q) {
        __syncthreads();
        fill_c(sC, e0, q);
        __syncthreads();
        s += sC[(tid * 7) % (NE * CSTR)];
    }
    if (e0 < nelem && tid < 120) out[(size_t)e0 * 360 + tid] = s;
}

// VALU: thread per (element, upper block)
__global__ __launch_bounds__(WG) void acc_valu(int nelem, double *out) {
    __shared__ double sC[NE * CSTR];
    const int tid = threadIdx.x, e0 = blockIdx.x * NE;
    const int le3 = tid / 10, blk = tid % 10;
    int ba = 0, bb = 0, cnt = 0;
    for (int a = 0; a < 4; ++a)
        for (int b = a; b < 4; ++b) {
            if (cnt == blk) { ba = a; bb = b; }
            ++cnt;
        }
    const bool cact = tid < NE * 10 && e0 + le3 < nelem;
    double acc[6][6];
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) acc[r][c] = 0.0;
    for (int q = 0; q < 4; ++q) {
        __syncthreads();
        fill_c(sC, e0, q);
        __syncthreads();
        if (cact) {
            const double *cbp = sC + le3 * CSTR;
            for (int s = 0; s < 8; ++s) {
                double va[6], vb[6];
                for (int r = 0; r < 6; ++r) {
                    va[r] = cbp[s * 24 + 6 * ba + r];
                    vb[r] = cbp[s * 24 + 6 * bb + r];
                }
                for (int r = 0; r < 6; ++r)
                    for (int c = 0; c < 6; ++c) acc[r][c] += va[r] * vb[c];
            }
        }
    }
    if (cact) {
        double *o = out + ((size_t)(e0 + le3) * 10 + blk) * 36;
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) o[r * 6 + c] = acc[r][c];
    }
}

__device__ __forceinline__ void tile_of(int T, int &ti, int &tj) {
    if (T > 20) T = 20;  // idle blocks of the last group repeat the last tile
    int c = 0;
    ti = tj = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 6; ++j) {
            if (c == T) { ti = i; tj = j; }
            ++c;
        }
}

// MFMA 4x4x4 (4 blocks), layout of gfx950: block b = (lane / 4) % 4;
// A[i][k]: i = lane % 4, k = lane / 16; B[k][j]: k = lane / 16, j = lane % 4;
// D[i][j]: i = lane / 16, j = lane % 4.  Wave w of the workgroup takes
// elements 6w..6w+5; group g, block b holds tile T = 4 g + b of the 21 upper
// 4x4 tiles (ti <= tj) of the 24 x 24 result: A[i][k] = C[k][4 ti + i],
// B[k][j] = C[k][4 tj + j].  The six groups' operands of one (element,
// k step) are loaded before its six MFMAs.
__global__ __launch_bounds__(WG) void acc_mfma(int nelem, double *out) {
    __shared__ double sC[NE * CSTR];
    const int tid = threadIdx.x, e0 = blockIdx.x * NE;
    const int w = tid / 64, l = tid % 64;
    const int b = (l >> 2) & 3, ij = l & 3, kk = l >> 4;
    int aoff[6], boff[6], dti[6], dtj[6];
    for (int g = 0; g < 6; ++g) {
        int ti, tj;
        tile_of(4 * g + b, ti, tj);
        aoff[g] = 4 * ti + ij;
        boff[g] = 4 * tj + ij;
        dti[g] = ti;
        dtj[g] = tj;
    }
    double acc[6][6];  // [element of the wave][group]
    for (int x = 0; x < 6; ++x)
        for (int g = 0; g < 6; ++g) acc[x][g] = 0.0;
    for (int q = 0; q < 4; ++q) {
        __syncthreads();
        fill_c(sC, e0, q);
        __syncthreads();
        for (int x = 0; x < 6; ++x) {
            const double *cbp = sC + (6 * w + x) * CSTR;
            for (int ks = 0; ks < 2; ++ks) {
                const double *row = cbp + (4 * ks + kk) * 24;
                double av[6], bv[6];
                for (int g = 0; g < 6; ++g) {
                    av[g] = row[aoff[g]];
                    bv[g] = row[boff[g]];
                }
                for (int g = 0; g < 6; ++g) acc[x][g] = __builtin_amdgcn_mfma_f64_4x4x4f64(av[g], bv[g], acc[x][g], 0, 0, 0);
            }
        }
    }
    // the upper 4x4 tiles into the node-block layout of acc_valu
    for (int x = 0; x < 6; ++x) {
        const int e = e0 + 6 * w + x;
        if (e >= nelem) continue;
        for (int g = 0; g < 6; ++g) {
            if (4 * g + b > 20) continue;
            const int i = 4 * dti[g] + kk, j = 4 * dtj[g] + ij;
            const int na = i / 6, nb = j / 6;
            if (na > nb) continue;  // below the node-block diagonal
            int bl = 0;
            for (int a2 = 0; a2 < na; ++a2) bl += 4 - a2;
            bl += nb - na;
            out[((size_t)e * 10 + bl) * 36 + (i % 6) * 6 + j % 6] = acc[x][g];
            if (na == nb) out[((size_t)e * 10 + bl) * 36 + (j % 6) * 6 + i % 6] = acc[x][g];
        }
    }
}

// Lane roles of an MFMA from the one-hot runs (out[L][p]: A one-hot at lane
// L, B[lane] = lane + 1, p an output position lane * nv + value): blocks are
// the connected components of output positions; within a block, A lanes
// with the same output positions share a row, B lanes seen at the same
// positions share a column, and the B lanes an A lane meets share its k.
struct Layout {
    int nblk = 0, m = 0, n = 0, k = 0;
    int a_blk[64], a_i[64], a_k[64], b_blk[64], b_k[64], b_j[64];
    int d_blk[256], d_i[256], d_j[256];
};
static bool deduce(const std::vector<double> &out, int npos, Layout &L) {
    std::vector<std::vector<int>> P(64), Q(64), meet(64);
    for (int a = 0; a < 64; ++a)
        for (int p = 0; p < npos; ++p) {
            const double v = out[(size_t)a * npos + p];
            if (v == 0.0) continue;
            const int bl = (int)v - 1;
            if (bl < 0 || bl > 63 || v != bl + 1.0) return false;
            P[a].push_back(p);
            Q[bl].push_back(p);
            meet[a].push_back(bl);
        }
    for (int b = 0; b < 64; ++b) {
        std::sort(Q[b].begin(), Q[b].end());
        Q[b].erase(std::unique(Q[b].begin(), Q[b].end()), Q[b].end());
    }
    for (int a = 0; a < 64; ++a) std::sort(meet[a].begin(), meet[a].end());
    std::vector<int> par(npos);
    for (int p = 0; p < npos; ++p) par[p] = p;
    auto find = [&](int x) {
        while (par[x] != x) x = par[x] = par[par[x]];
        return x;
    };
    auto unite = [&](const std::vector<int> &v) {
        for (size_t t = 1; t < v.size(); ++t) par[find(v[t])] = find(v[0]);
    };
    for (int a = 0; a < 64; ++a) unite(P[a]);
    for (int b = 0; b < 64; ++b) unite(Q[b]);
    std::map<int, int> blk;
    for (int p = 0; p < npos; ++p)
        if (!blk.count(find(p))) {
            const int id = (int)blk.size();
            blk[find(p)] = id;
        }
    L.nblk = (int)blk.size();
    std::vector<std::map<std::vector<int>, int>> rows(L.nblk), cols(L.nblk), ks(L.nblk);
    for (int a = 0; a < 64; ++a) {
        if (P[a].empty()) return false;
        const int bk = blk[find(P[a][0])];
        L.a_blk[a] = bk;
        auto &R = rows[bk];
        if (!R.count(P[a])) { const int id = (int)R.size(); R[P[a]] = id; }
        L.a_i[a] = R[P[a]];
        auto &K = ks[bk];
        if (!K.count(meet[a])) { const int id = (int)K.size(); K[meet[a]] = id; }
        L.a_k[a] = K[meet[a]];
    }
    for (int b = 0; b < 64; ++b) {
        if (Q[b].empty()) return false;
        const int bk = blk[find(Q[b][0])];
        L.b_blk[b] = bk;
        auto &Cc = cols[bk];
        if (!Cc.count(Q[b])) { const int id = (int)Cc.size(); Cc[Q[b]] = id; }
        L.b_j[b] = Cc[Q[b]];
        L.b_k[b] = -1;
        for (auto &kv : ks[bk])
            if (std::binary_search(kv.first.begin(), kv.first.end(), b)) L.b_k[b] = kv.second;
        if (L.b_k[b] < 0) return false;
    }
    for (int p = 0; p < npos; ++p) L.d_blk[p] = -1;
    for (int bk = 0; bk < L.nblk; ++bk) {
        for (auto &kv : rows[bk])
            for (int p : kv.first) { L.d_blk[p] = bk; L.d_i[p] = kv.second; }
        for (auto &kv : cols[bk])
            for (int p : kv.first) L.d_j[p] = kv.second;
    }
    for (int p = 0; p < npos; ++p)
        if (L.d_blk[p] < 0) return false;
    L.m = (int)rows[0].size();
    L.n = (int)cols[0].size();
    L.k = (int)ks[0].size();
    return true;
}

// The deduced 4x4x4 roles against the formula acc_mfma uses: the labels of
// blocks, rows, columns and k are each fixed up to a relabelling, so the
// check compares the partitions (lanes with equal formula values must have
// equal deduced labels and the other way round).
static bool same_partition(const std::vector<int> &x, const std::vector<int> &y) {
    std::map<int, int> f, g;
    for (size_t t = 0; t < x.size(); ++t) {
        if (f.count(x[t]) && f[x[t]] != y[t]) return false;
        if (g.count(y[t]) && g[y[t]] != x[t]) return false;
        f[x[t]] = y[t];
        g[y[t]] = x[t];
    }
    return true;
}
static bool layout4_matches(const Layout &L) {
    std::vector<int> ab, ai, ak, bb, bk, bj, db, di, dj, fb, fi, fk, fj, fdi;
    for (int l = 0; l < 64; ++l) {
        ab.push_back(L.a_blk[l]); ai.push_back(L.a_blk[l] * 4 + L.a_i[l]); ak.push_back(L.a_blk[l] * 4 + L.a_k[l]);
        bb.push_back(L.b_blk[l]); bk.push_back(L.b_blk[l] * 4 + L.b_k[l]); bj.push_back(L.b_blk[l] * 4 + L.b_j[l]);
        db.push_back(L.d_blk[l]); di.push_back(L.d_blk[l] * 4 + L.d_i[l]); dj.push_back(L.d_blk[l] * 4 + L.d_j[l]);
        const int b = (l >> 2) & 3;
        fb.push_back(b); fi.push_back(b * 4 + (l & 3)); fk.push_back(b * 4 + (l >> 4)); fj.push_back(b * 4 + (l & 3));
        fdi.push_back(b * 4 + (l >> 4));
    }
    // A: (block, i = lane % 4, k = lane / 16); B: (block, k = lane / 16,
    // j = lane % 4); D: (block, i = lane / 16, j = lane % 4); and A's k
    // must be B's k where both are lane / 16, A's i and D's j both lane % 4
    if (!(same_partition(ab, fb) && same_partition(ai, fi) && same_partition(ak, fk) && same_partition(bb, fb) &&
          same_partition(bk, fk) && same_partition(bj, fj) && same_partition(db, fb) && same_partition(di, fdi) &&
          same_partition(dj, fj)))
        return false;
    // across roles, within a block: A's k is B's k, A's row is D's row, B's
    // column is D's column
    for (int l = 0; l < 64; ++l)
        for (int m = 0; m < 64; ++m) {
            if (((l >> 2) & 3) != ((m >> 2) & 3)) continue;
            if (L.a_blk[l] != L.b_blk[m] || L.a_blk[l] != L.d_blk[m]) return false;
            if (((l >> 4) == (m >> 4)) != (L.a_k[l] == L.b_k[m])) return false;
            if (((l & 3) == (m >> 4)) != (L.a_i[l] == L.d_i[m])) return false;
            if (((l & 3) == (m & 3)) != (L.b_j[l] == L.d_j[m])) return false;
        }
    return true;
}

int main() {
    double *dO;
    CK(hipMalloc(&dO, 64 * 256 * 8));
    std::vector<double> O(64 * 256);
    hipLaunchKernelGGL(onehot16, dim3(64), dim3(64), 0, 0, dO);
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    CK(hipMemcpy(O.data(), dO, 64 * 256 * 8, hipMemcpyDeviceToHost));
    Layout L16, L4;
    const bool ok16 = deduce(O, 256, L16);
    std::printf("v_mfma_f64_16x16x4f64: layout deduced %s (%d block, %d x %d, k %d)\n", ok16 ? "yes" : "no", L16.nblk,
                L16.m, L16.n, L16.k);
    hipLaunchKernelGGL(onehot4, dim3(64), dim3(64), 0, 0, dO);
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    CK(hipMemcpy(O.data(), dO, 64 * 64 * 8, hipMemcpyDeviceToHost));
    const bool ok4 = deduce(O, 64, L4);
    const bool match4 = ok4 && L4.nblk == 4 && L4.m == 4 && L4.n == 4 && L4.k == 4 && layout4_matches(L4);
    std::printf("v_mfma_f64_4x4x4f64: layout deduced %s (%d blocks, %d x %d, k %d); the accumulation kernel's "
                "layout %s\n", ok4 ? "yes" : "no", L4.nblk, L4.m, L4.n, L4.k, match4 ? "matches" : "DOES NOT MATCH");
    CK(hipFree(dO));

    hipEvent_t t0, t1;
    CK(hipEventCreate(&t0));
    CK(hipEventCreate(&t1));
    double *dT;
    CK(hipMalloc(&dT, (size_t)256 * 4 * 2 * 64 * 8));
    const int iters = 4000;
    const char *names[3] = {"v_mfma_f64_16x16x4f64", "v_mfma_f64_4x4x4f64 (4 blocks)", "v_fma_f64 (8 chains)"};
    for (int kind = 0; kind < 3; ++kind)
        for (int wps = 1; wps <= 2; ++wps) {
            const int blocks = 256 * 4 * wps;  // waves per SIMD on 256 CUs x 4 SIMDs
            auto launch = [&]() {
                if (kind == 0) hipLaunchKernelGGL(issue<0>, dim3(blocks), dim3(64), 0, 0, iters, dT);
                else if (kind == 1) hipLaunchKernelGGL(issue<1>, dim3(blocks), dim3(64), 0, 0, iters, dT);
                else hipLaunchKernelGGL(issue<2>, dim3(blocks), dim3(64), 0, 0, iters, dT);
            };
            launch();
            CK(hipGetLastError());
            CK(hipDeviceSynchronize());
            CK(hipEventRecord(t0));
            launch();
            CK(hipEventRecord(t1));
            CK(hipEventSynchronize(t1));
            float ms;
            CK(hipEventElapsedTime(&ms, t0, t1));
            const double per_simd = (double)blocks * iters * (kind == 2 ? 8 : 4) / 1024.0;
            std::printf("%-32s %d wave(s)/SIMD: %.1f cycles per wave instruction per SIMD at 2.4 GHz\n", names[kind],
                        wps, ms * 1e6 / per_simd * 2.4);
        }
    CK(hipFree(dT));

    const int nel = 262144;
    double *dK1, *dK2, *dF;
    CK(hipMalloc(&dK1, (size_t)nel * 360 * 8));
    CK(hipMalloc(&dK2, (size_t)nel * 360 * 8));
    CK(hipMalloc(&dF, (size_t)nel * 360 * 8));
    const int nb = (nel + NE - 1) / NE;
    auto timed = [&](void (*k)(int, double *), double *dst) {
        hipLaunchKernelGGL(k, dim3(nb), dim3(WG), 0, 0, nel, dst);
        CK(hipGetLastError());
        float best = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
            CK(hipEventRecord(t0));
            hipLaunchKernelGGL(k, dim3(nb), dim3(WG), 0, 0, nel, dst);
            CK(hipEventRecord(t1));
            CK(hipEventSynchronize(t1));
            float ms;
            CK(hipEventElapsedTime(&ms, t0, t1));
            best = std::min(best, ms);
        }
        return best;
    };
    const float tf = timed(acc_fill, dF), tv = timed(acc_valu, dK1), tm = timed(acc_mfma, dK2);
    std::printf("accumulation, 262144 Q4 elements: fill of C %.3f ms; VALU %.3f ms (%.3f without the fill); "
                "MFMA 4x4x4 %.3f ms (%.3f without the fill)\n", tf, tv, tv - tf, tm, tm - tf);
    std::vector<double> K1((size_t)nel * 360), K2((size_t)nel * 360);
    CK(hipMemcpy(K1.data(), dK1, K1.size() * 8, hipMemcpyDeviceToHost));
    CK(hipMemcpy(K2.data(), dK2, K2.size() * 8, hipMemcpyDeviceToHost));
    size_t diff = 0;
    for (size_t i = 0; i < K1.size(); ++i) diff += K1[i] != K2[i];
    std::printf("MFMA blocks equal to the VALU blocks: %s (%zu of %zu values differ)\n", diff == 0 ? "yes" : "NO", diff,
                K1.size());
    return (ok16 && match4 && diff == 0) ? 0 : 1;
}
