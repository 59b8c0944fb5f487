// 3D Laguerre-cell pass.  This file holds the thread-per-cell kernel: each
// thread owns one polyhedron slot in HBM (element k of slot s at k * S + s,
// so a wavefront's accesses coalesce) and runs the per-cell routine of
// cell3d_impl.h over a grid-stride range of cells in grid order.
#include "cell3d_impl.h"

namespace sdot_amd {

namespace {

// capacity tiers of the thread kernel: cells that outgrow the first tier (or the
// 64-lane wavefront kernel) are re-run on the second, whose plane indices still
// fit the 8-bit fields of the vertex plane triples
constexpr int VCAP3 = 160, PCAP3 = 96;      // tier 1: vertices / planes per polyhedron
constexpr int VCAP3W = 508, PCAP3W = 255;   // tier 2
constexpr int BS3 = 256;

template <class T>
size_t slot_bytes(int vcap, int pcap, bool jac = false) {
    return (size_t)vcap * (7 * sizeof(T) + 2 * sizeof(int)) + (size_t)pcap * (4 * sizeof(T) + sizeof(int) + 2 * sizeof(double)) +
           (size_t)pcap * pcap * sizeof(unsigned short) + (jac ? (size_t)pcap * 19 * sizeof(double) : 0);
}

template <class T>
__host__ __device__ c3::Poly3Mem<T> slot(char *base, int64_t S, int64_t s, int vcap, int pcap, bool jac = false) {
    // field-major layout: [field][k][slot]
    c3::Poly3Mem<T> m;
    char *p = base;
    auto take = [&](auto *&dst, size_t elem, int cap) {
        using P = typename std::remove_reference<decltype(dst)>::type;
        dst = (P)(p) + s;
        p += elem * cap * S;
    };
    take(m.px, sizeof(T), vcap);
    take(m.py, sizeof(T), vcap);
    take(m.pz, sizeof(T), vcap);
    take(m.d, sizeof(T), vcap);
    take(m.sx, sizeof(T), vcap);
    take(m.sy, sizeof(T), vcap);
    take(m.sz, sizeof(T), vcap);
    take(m.tri, sizeof(int), vcap);
    take(m.stri, sizeof(int), vcap);
    take(m.nx, sizeof(T), pcap);
    take(m.ny, sizeof(T), pcap);
    take(m.nz, sizeof(T), pcap);
    take(m.nh, sizeof(T), pcap);
    take(m.pid, sizeof(int), pcap);
    take(m.acc, sizeof(double), pcap);
    take(m.hacc, sizeof(double), pcap);
    take(m.emap, sizeof(unsigned short), pcap * pcap);
    if (jac) {
        take(m.jmom, sizeof(double), pcap * 9);
        take(m.jsum, sizeof(double), pcap * 10);
    }
    m.S = (int)S;
    return m;
}

// LAST: overflowing cells raise FLAG_CELL_OVERFLOW; otherwise they are appended
// to ovf_list (counter ovf_count) for the next tier
template <class T, int VC, int PC, bool LAST, bool JAC = false>
__global__ __launch_bounds__(BS3) void k_cells3_thread(const CellArgs *__restrict__ ap, char *scratch, int64_t nslots) {
    const CellArgs &a = *ap;
    const int64_t s = blockIdx.x * (int64_t)BS3 + threadIdx.x;
    if (s >= nslots) return;
    const c3::Poly3Mem<T> m = slot<T>(scratch, nslots, s, VC, PC, JAC);
    const int64_t count = a.work ? a.work_n : a.N;
    for (int64_t k = s; k < count; k += nslots) {
        const int64_t t = a.work ? a.work[k] : k;
        if (c3::cell3_body<T>(a, t, m, VC, PC)) {
            if (LAST) {
                flag_or(a.flags, FLAG_CELL_OVERFLOW);
            } else {
                const unsigned q = atomicAdd(a.ovf_count, 1u);
                a.ovf_list[q] = (int)t;
            }
        }
    }
}

// the arguments go to HBM and the kernels read them through a pointer: a by-value
// argument block whose address is taken would be copied to scratch per lane
const CellArgs *upload_args(sdot_pd *pd, const CellArgs &a, int slot) {
    pd->dargs.ensure(2);
    pd->hargs = a;
    HIPCK(hipMemcpyAsync(pd->dargs.p + slot, &pd->hargs, sizeof(CellArgs), hipMemcpyHostToDevice, pd->stream));
    HIPCK(hipStreamSynchronize(pd->stream));
    return pd->dargs.p + slot;
}

template <class T, int VC, int PC, bool LAST, bool JAC = false>
void launch_thread3(sdot_pd *pd, const CellArgs &a) {
    int dev = 0, ncu = 0;
    HIPCK(hipGetDevice(&dev));
    HIPCK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int64_t count = a.work ? a.work_n : a.N;
    int64_t nslots = (int64_t)ncu * 4 * 64;     // four wavefronts per CU
    if (nslots > count) nslots = count;
    if (nslots < 1) return;
    pd->scratch3.ensure(slot_bytes<T>(VC, PC, JAC) * nslots);
    const unsigned nb = (unsigned)((nslots + BS3 - 1) / BS3);
    const CellArgs *da = upload_args(pd, a, 1);
    hipLaunchKernelGGL((k_cells3_thread<T, VC, PC, LAST, JAC>), dim3(nb), dim3(BS3), 0, pd->stream, da, pd->scratch3.p,
                       nslots);
    HIPCK(hipGetLastError());
}


// ---------------------------------------------------------------------------
// Wavefront-cooperative kernel: one cell per wavefront, one polyhedron vertex
// per lane (position and plane triple in registers), planes and the dual-edge
// map in the wavefront's LDS slice.  A candidate costs one dot product per
// lane and a ballot; a cut builds the new vertices on the crossed edges with
// lane shuffles, compacts the vertex list with a wavefront prefix sum and
// rewrites the edge map.  Cells that need more than 64 vertices or planes go
// to the thread kernel above.
// ---------------------------------------------------------------------------
constexpr int WV = 64;            // lanes = vertex capacity = plane capacity
constexpr int FRONT = 192;        // far-field traversal: frontier capacity (nodes)
constexpr int NEAR_RINGS = 1;     // rings of grid cells visited before the far-field traversal
constexpr int WPB = 4;            // wavefronts per block

struct WaveLds {
    double pn[WV][4];             // planes (FP64): n.q <= h
    int pid[WV];
    unsigned char edge[WV * WV];  // edge[x * 64 + y] = lane holding the dual edge x -> y
    double acc[WV];               // face-area accumulators
    double hacc[WV];              // density-weighted face areas (image mode)
    double sx[WV], sy[WV], sz[WV];
    int stri[WV];
    int cpre[WV], ccs[WV], ccnt[WV];   // staged grid cells: candidate prefix, first dirac, count
    uint32_t fr[2][FRONT];        // far-field traversal: node frontier (double buffered)
};

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}
__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ inline int wave_sum_i(int v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
// Cross-lane helpers without LDS traffic (the shuffles above go through the LDS
// crossbar, one dependent round trip per step).
// Value of v in lane k, k uniform across the wavefront (v_readlane).
__device__ inline int lane_value(int v, int k) { return __builtin_amdgcn_readlane(v, k); }
__device__ inline float lane_value(float v, int k) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k)); }
__device__ inline double lane_value(double v, int k) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// DPP move within rows of 16 lanes; CTRL is a full quad permutation or a row
// rotation, so every lane has a source lane
template <int CTRL>
__device__ inline double dpp_row(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
constexpr int DPP_QUAD_SWAP1 = 0xB1;   // quad_perm [1,0,3,2]
constexpr int DPP_QUAD_SWAP2 = 0x4E;   // quad_perm [2,3,0,1]
constexpr int DPP_ROW_ROR4 = 0x124, DPP_ROW_ROR8 = 0x128;
// reductions over the whole wavefront (every lane active): within each row by
// quad swaps and row rotations (after them every lane holds its row's result),
// then the four rows from lanes 0, 16, 32, 48
__device__ inline double wave_max_dpp(double v) {
    v = fmax(v, dpp_row<DPP_QUAD_SWAP1>(v));
    v = fmax(v, dpp_row<DPP_QUAD_SWAP2>(v));
    v = fmax(v, dpp_row<DPP_ROW_ROR4>(v));
    v = fmax(v, dpp_row<DPP_ROW_ROR8>(v));
    return fmax(fmax(lane_value(v, 0), lane_value(v, 16)), fmax(lane_value(v, 32), lane_value(v, 48)));
}
__device__ inline double wave_sum_dpp(double v) {
    v += dpp_row<DPP_QUAD_SWAP1>(v);
    v += dpp_row<DPP_QUAD_SWAP2>(v);
    v += dpp_row<DPP_ROW_ROR4>(v);
    v += dpp_row<DPP_ROW_ROR8>(v);
    return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}
// exclusive prefix sum of small values (0 <= v < 8) from three ballots
__device__ inline int wave_excl_scan_small(int v) {
    int r = 0;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const unsigned long long m = __ballot((v >> b) & 1);
        r += (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u)) << b;
    }
    return r;
}

__device__ inline int wave_excl_scan(int v, int lane) {
    int x = v;
    for (int o = 1; o < WV; o <<= 1) {
        int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    return x - v;
}

template <class T>
struct WaveCell {
    const CellArgs &a;
    WaveLds &L;
    const int lane;
    int nv = 0, np = 0;
    // this lane's vertex
    T x = 0, y = 0, z = 0;
    int tri = 0;
    double c0x, c0y, c0z, w0;
    // current voxel box (local coordinates), kept as scalars so the object stays in registers
    double vlo0 = 0, vlo1 = 0, vlo2 = 0, vhi0 = 0, vhi1 = 0, vhi2 = 0;
    __device__ double vlo(int axis) const { return axis == 0 ? vlo0 : (axis == 1 ? vlo1 : vlo2); }
    __device__ double vhi(int axis) const { return axis == 0 ? vhi0 : (axis == 1 ? vhi1 : vhi2); }

    __device__ WaveCell(const CellArgs &a_, WaveLds &L_, int lane_) : a(a_), L(L_), lane(lane_) {}

    __device__ void set_plane(int k, double nx, double ny, double nz, double h, int id) {
        if (lane == 0) {
            L.pn[k][0] = nx;
            L.pn[k][1] = ny;
            L.pn[k][2] = nz;
            L.pn[k][3] = h;
            L.pid[k] = id;
        }
    }

    __device__ void write_edges() {
        if (lane < nv) {
            const int p0 = tri & 255, p1 = (tri >> 8) & 255, p2 = (tri >> 16) & 255;
            L.edge[p0 * WV + p1] = (unsigned char)lane;
            L.edge[p1 * WV + p2] = (unsigned char)lane;
            L.edge[p2 * WV + p0] = (unsigned char)lane;
        }
        wave_sync();
    }

    __device__ void init_box() {
        if (lane < 6) {
            // (selects, not an indexed array: a per-lane index would put the array in memory)
            const int axis = lane >> 1, up = lane & 1;
            const double c0a = axis == 0 ? c0x : (axis == 1 ? c0y : c0z);
            const double lo = axis == 0 ? a.dom_lo[0] : (axis == 1 ? a.dom_lo[1] : a.dom_lo[2]);
            const double hi = axis == 0 ? a.dom_hi[0] : (axis == 1 ? a.dom_hi[1] : a.dom_hi[2]);
            const double off = (up ? hi : lo) - c0a;
            L.pn[lane][0] = axis == 0 ? (up ? 1.0 : -1.0) : 0.0;
            L.pn[lane][1] = axis == 1 ? (up ? 1.0 : -1.0) : 0.0;
            L.pn[lane][2] = axis == 2 ? (up ? 1.0 : -1.0) : 0.0;
            L.pn[lane][3] = up ? off : -off;
            L.pid[lane] = -1 - lane;
        }
        np = 6;
        nv = 8;
        if (lane < 8) {
            const int sx = lane & 1, sy = (lane >> 1) & 1, sz = (lane >> 2) & 1;
            int pa = sx, pb = 2 + sy, pc = 4 + sz;
            if ((sx ? 1 : -1) * (sy ? 1 : -1) * (sz ? 1 : -1) < 0) {
                int t = pb;
                pb = pc;
                pc = t;
            }
            tri = pa | (pb << 8) | (pc << 16);
            x = T((sx ? a.dom_hi[0] : a.dom_lo[0]) - c0x);
            y = T((sy ? a.dom_hi[1] : a.dom_lo[1]) - c0y);
            z = T((sz ? a.dom_hi[2] : a.dom_lo[2]) - c0z);
        }
        write_edges();
    }

    __device__ double radius2() const {
        double r = lane < nv ? (double)x * x + (double)y * y + (double)z * z : 0.0;
        return wave_max_dpp(r);
    }

    // a sphere holding the polyhedron: centre the mean of the vertices, squared radius
    __device__ void sphere(double &mx, double &my, double &mz, double &r2) const {
        const bool v = lane < nv;
        const double inv = 1.0 / nv;
        mx = wave_sum_dpp(v ? (double)x : 0.0) * inv;
        my = wave_sum_dpp(v ? (double)y : 0.0) * inv;
        mz = wave_sum_dpp(v ? (double)z : 0.0) * inv;
        const double dx = (double)x - mx, dy = (double)y - my, dz = (double)z - mz;
        r2 = wave_max_dpp(v ? dx * dx + dy * dy + dz * dz : 0.0);
    }

    // Side of this lane's vertex for the plane n.q <= h, exactly from the planes:
    // with rows [n_k, h_k] of the vertex's planes a, b, c and the new plane p,
    // det4 = (h_p - n_p.v) det3 and det3 = det(n_a, n_b, n_c) > 0 (orientation).
    __device__ double tie_distance(double nx, double ny, double nz, double h) const {
        double m[4][4];
        plane64(tri & 255, m[0][0], m[0][1], m[0][2], m[0][3]);
        plane64((tri >> 8) & 255, m[1][0], m[1][1], m[1][2], m[1][3]);
        plane64((tri >> 16) & 255, m[2][0], m[2][1], m[2][2], m[2][3]);
        m[3][0] = nx;
        m[3][1] = ny;
        m[3][2] = nz;
        m[3][3] = h;
        const double s0 = m[0][0] * m[1][1] - m[1][0] * m[0][1], s1 = m[0][0] * m[1][2] - m[1][0] * m[0][2];
        const double s2 = m[0][0] * m[1][3] - m[1][0] * m[0][3], s3 = m[0][1] * m[1][2] - m[1][1] * m[0][2];
        const double s4 = m[0][1] * m[1][3] - m[1][1] * m[0][3], s5 = m[0][2] * m[1][3] - m[1][2] * m[0][3];
        const double c5 = m[2][2] * m[3][3] - m[3][2] * m[2][3], c4 = m[2][1] * m[3][3] - m[3][1] * m[2][3];
        const double c3 = m[2][1] * m[3][2] - m[3][1] * m[2][2], c2 = m[2][0] * m[3][3] - m[3][0] * m[2][3];
        const double c1 = m[2][0] * m[3][2] - m[3][0] * m[2][2], c0 = m[2][0] * m[3][1] - m[3][0] * m[2][1];
        const double det4 = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
        const double det3 = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                            m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
        return det3 > 0 ? det4 / det3 : (double)(h - (nx * x + ny * y + nz * z));
    }

    // renumber the planes still bounding a face (slots of removed planes are freed)
    __device__ void compact_planes() {
        const unsigned long long used = used_planes();
        const bool keep = lane < np && ((used >> lane) & 1ull);
        double q0 = 0, q1 = 0, q2 = 0, q3 = 0;
        int qid = 0;
        if (keep) {
            q0 = L.pn[lane][0];
            q1 = L.pn[lane][1];
            q2 = L.pn[lane][2];
            q3 = L.pn[lane][3];
            qid = L.pid[lane];
        }
        wave_sync();
        const int dst = __popcll(used & ((1ull << lane) - 1ull));
        if (keep) {
            L.pn[dst][0] = q0;
            L.pn[dst][1] = q1;
            L.pn[dst][2] = q2;
            L.pn[dst][3] = q3;
            L.pid[dst] = qid;
        }
        if (lane < nv) {
            const int a0 = tri & 255, a1 = (tri >> 8) & 255, a2 = (tri >> 16) & 255;
            const int b0 = __popcll(used & ((1ull << a0) - 1ull)), b1 = __popcll(used & ((1ull << a1) - 1ull)),
                      b2 = __popcll(used & ((1ull << a2) - 1ull));
            tri = b0 | (b1 << 8) | (b2 << 16);
        }
        np = __popcll(used);
        wave_sync();
        write_edges();
    }

    // cut by q.n <= h: 0 no change, 1 cut (nv may be 0), 2 capacity exceeded
    // always inlined: an out-of-line call saves the wavefront's live registers to scratch
    // around every cut (the compiler chose that for the warm-started instantiation, which
    // ran twice as slow)
    __device__ __attribute__((always_inline)) int cut(T nx, T ny, T nz, T h, int id, double nx64, double ny64,
                                                      double nz64, double h64, double tie_tol, bool may_compact) {
        const bool valid = lane < nv;
        T d = valid ? fma(-x, nx, fma(-y, ny, fma(-z, nz, h))) : T(1);
        bool out = valid && d < 0;
        const bool near = valid && fabs((double)d) <= tie_tol;
        if (__ballot(near)) {
            if (near) {
                const double dd = tie_distance(nx64, ny64, nz64, h64);
                out = dd < 0;
                d = T(dd);
                if (out && !(d < 0)) d = -T(1e-30);
            }
        }
        const unsigned long long mo = __ballot(out);
        if (!mo) return 0;
        if (__popcll(mo) == nv) {
            nv = 0;
            return 1;
        }
        if (np >= WV && may_compact) compact_planes();
        if (np >= WV) return 2;
        const int pn = np;
        const int p[3] = {tri & 255, (tri >> 8) & 255, (tri >> 16) & 255};
        int cnt = out ? 0 : (valid ? 1 : 0);
        T nxv[3], nyv[3], nzv[3];
        int ntri[3];
        bool mk[3];
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const int xa = p[e], yb = p[e == 2 ? 0 : e + 1];
            const int nb = out ? (int)L.edge[yb * WV + xa] : lane;
            const T dn = __shfl(d, nb), xn = __shfl(x, nb), yn = __shfl(y, nb), zn = __shfl(z, nb);
            mk[e] = out && !(dn < 0);
            const T mm = dn / (d - dn);
            nxv[e] = xn - mm * (x - xn);
            nyv[e] = yn - mm * (y - yn);
            nzv[e] = zn - mm * (z - zn);
            ntri[e] = xa | (yb << 8) | (pn << 16);
            cnt += mk[e] ? 1 : 0;
        }
        const int pre = wave_excl_scan_small(cnt);   // cnt <= 3
        const int total = lane_value(pre + cnt, WV - 1);
        if (total > WV) return 2;
        wave_sync();
        if (valid && !out) {
            L.sx[pre] = x;
            L.sy[pre] = y;
            L.sz[pre] = z;
            L.stri[pre] = tri;
        }
        int o = pre;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            if (mk[e]) {
                L.sx[o] = nxv[e];
                L.sy[o] = nyv[e];
                L.sz[o] = nzv[e];
                L.stri[o] = ntri[e];
                ++o;
            }
        }
        set_plane(pn, nx64, ny64, nz64, h64, id);
        wave_sync();
        nv = total;
        np = pn + 1;
        if (lane < nv) {
            x = T(L.sx[lane]);
            y = T(L.sy[lane]);
            z = T(L.sz[lane]);
            tri = L.stri[lane];
        }
        wave_sync();
        write_edges();
        return 1;
    }

    // every plane of the polyhedron, voxel sides included, is in the LDS table (set_plane
    // at its cut); reading it there keeps the voxel bounds out of a runtime-indexed
    // selection, which the compiler turned into an indexed load of the whole cell state
    // from scratch memory
    __device__ void plane64(int k, double &nx, double &ny, double &nz, double &h) const {
        nx = L.pn[k][0];
        ny = L.pn[k][1];
        nz = L.pn[k][2];
        h = L.pn[k][3];
    }

    // FP64 position of this lane's vertex (intersection of its planes in the mixed path)
    __device__ void vertex64(double &X, double &Y, double &Z) const {
        X = x;
        Y = y;
        Z = z;
        if (sizeof(T) == sizeof(double) || lane >= nv) return;
        double n0[4], n1[4], n2[4];
        plane64(tri & 255, n0[0], n0[1], n0[2], n0[3]);
        plane64((tri >> 8) & 255, n1[0], n1[1], n1[2], n1[3]);
        plane64((tri >> 16) & 255, n2[0], n2[1], n2[2], n2[3]);
        const double c12x = n1[1] * n2[2] - n1[2] * n2[1], c12y = n1[2] * n2[0] - n1[0] * n2[2], c12z = n1[0] * n2[1] - n1[1] * n2[0];
        const double c20x = n2[1] * n0[2] - n2[2] * n0[1], c20y = n2[2] * n0[0] - n2[0] * n0[2], c20z = n2[0] * n0[1] - n2[1] * n0[0];
        const double c01x = n0[1] * n1[2] - n0[2] * n1[1], c01y = n0[2] * n1[0] - n0[0] * n1[2], c01z = n0[0] * n1[1] - n0[1] * n1[0];
        const double det = n0[0] * c12x + n0[1] * c12y + n0[2] * c12z;
        const double nrm = sqrt((n0[0] * n0[0] + n0[1] * n0[1] + n0[2] * n0[2]) * (n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2]) *
                                (n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2]));
        if (!(fabs(det) > 1e-9 * nrm)) return;
        X = (n0[3] * c12x + n1[3] * c20x + n2[3] * c01x) / det;
        Y = (n0[3] * c12y + n1[3] * c20y + n2[3] * c01y) / det;
        Z = (n0[3] * c12z + n1[3] * c20z + n2[3] * c01z) / det;
    }

    // face areas into L.acc[0..np), returns the volume (all lanes).  With `mom`, also
    // the first moments about the dirac (all lanes): signed tetrahedra (dirac, foot
    // of the face plane, edge) summed over the directed edges of every face, scaled
    // to the volume (add_centroid_contrib, ConvexPolyhedron3.tcc:884-924)
    __device__ double integrate(double *mom = nullptr) {
        if (lane < np) L.acc[lane] = 0;
        wave_sync();
        double X, Y, Z;
        vertex64(X, Y, Z);
        const bool valid = lane < nv;
        const int p[3] = {tri & 255, (tri >> 8) & 255, (tri >> 16) & 255};
        double tv = 0, tmx = 0, tmy = 0, tmz = 0;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
            const int f = p[e], gq = p[e == 2 ? 0 : e + 1];
            const int q = valid ? (int)L.edge[gq * WV + f] : lane;
            const double Xq = __shfl(X, q), Yq = __shfl(Y, q), Zq = __shfl(Z, q);
            if (valid) {
                double fx, fy, fz, fh;
                plane64(f, fx, fy, fz, fh);
                const double cx = Y * Zq - Z * Yq, cy = Z * Xq - X * Zq, cz = X * Yq - Y * Xq;
                const double cn = cx * fx + cy * fy + cz * fz, n2 = fx * fx + fy * fy + fz * fz;
                atomicAdd(&L.acc[f], cn / sqrt(n2));
                if (mom) {
                    const double s = fh / n2, v = s * cn / 6;   // foot F = s n, volume F.(P x Q)/6
                    tv += v;
                    tmx += v * (s * fx + X + Xq) / 4;
                    tmy += v * (s * fy + Y + Yq) / 4;
                    tmz += v * (s * fz + Z + Zq) / 4;
                }
            }
        }
        if (mom) {
            const double sv = wave_sum_dpp(tv);
            const double k = sv != 0 ? 1.0 / sv : 0.0;   // moments / signed volume = centroid offset
            mom[0] = wave_sum_dpp(tmx) * k;
            mom[1] = wave_sum_dpp(tmy) * k;
            mom[2] = wave_sum_dpp(tmz) * k;
        }
        wave_sync();
        double vol = 0;
        if (lane < np) {
            const double ar = 0.5 * fabs(L.acc[lane]);
            L.acc[lane] = ar;
            double fx, fy, fz, fh;
            plane64(lane, fx, fy, fz, fh);
            vol = fh / sqrt(fx * fx + fy * fy + fz * fz) * ar / 3;
        }
        wave_sync();
        return wave_sum_dpp(vol);
    }

    __device__ unsigned long long used_planes() const {
        unsigned long long m = 0;
        if (lane < nv) m = (1ull << (tri & 255)) | (1ull << ((tri >> 8) & 255)) | (1ull << ((tri >> 16) & 255));
        for (int o = 32; o > 0; o >>= 1) m |= __shfl_xor(m, o);
        return m;
    }
};

// WARM: the instantiation for warm-started passes (CellArgs::hint_col set); the cold
// one carries none of their code.  Three wavefronts per SIMD: registers capped at 168,
// LDS 3 x 4 x 11.5 KB per CU
template <class T, bool WARM>
__global__ __launch_bounds__(WV * WPB, 3) void k_cells3_wave(const CellArgs *__restrict__ ap) {
    const CellArgs &a = *ap;
    __shared__ WaveLds lds[WPB];
    const int lane = threadIdx.x & (WV - 1);
    const int wid = threadIdx.x / WV;
    WaveLds &L = lds[wid];
    const int64_t nwaves = (int64_t)gridDim.x * WPB;
    const int64_t w0id = __builtin_amdgcn_readfirstlane(xcd_tile(blockIdx.x, gridDim.x) * WPB + wid);
    const GridInfo &g = a.g;
    const double W = dbl_unorder(*a.wglob);
    int maxr = g.ncell[0];
    if (g.ncell[1] > maxr) maxr = g.ncell[1];
    if (g.ncell[2] > maxr) maxr = g.ncell[2];
    // cells in grid order, one contiguous range per wavefront
    const int64_t per = (a.N + nwaves - 1) / nwaves;
    const int64_t tb = w0id * per, te = tb + per < a.N ? tb + per : a.N;
    for (int64_t t = tb; t < te; ++t) {
        // stop_if_void: one read per cell, made uniform across the wavefront
        if (a.stop_if_void && __builtin_amdgcn_readfirstlane((int)void_seen(a, 0))) break;
        WaveCell<T> C(a, L, lane);
        C.c0x = a.spos[3 * t];
        C.c0y = a.spos[3 * t + 1];
        C.c0z = a.spos[3 * t + 2];
        C.w0 = a.sw[t];
        const double c0x = C.c0x, c0y = C.c0y, c0z = C.c0z, w0 = C.w0;
        const double c0[3] = {c0x, c0y, c0z};
        C.init_box();
        bool overflow = false, bad = false;
        if (a.nitems == 1) {   // one domain item: the starting cell is its box cut by its planes
            const double r0 = sqrt(C.radius2());
            for (int q = a.it_off[0]; q < a.it_off[1] && C.nv > 0 && !overflow; ++q) {
                const double *pl = a.it_planes + 4 * q;
                const double hl = pl[3] - (pl[0] * c0x + pl[1] * c0y + pl[2] * c0z);
                const double nn = sqrt(pl[0] * pl[0] + pl[1] * pl[1] + pl[2] * pl[2]);
                if (C.cut(T(pl[0]), T(pl[1]), T(pl[2]), T(hl), -100 - q, pl[0], pl[1], pl[2], hl,
                          c3::Eps3<T>::tie * nn * r0, true) == 2)
                    overflow = true;
            }
        }
        double R2 = C.radius2();
        const int hh[3] = {grid_coord(g, 0, c0x), grid_coord(g, 1, c0y), grid_coord(g, 2, c0z)};
        // Warm start (CellArgs::hint_col): this lane's neighbour of the cell in the last
        // complete pass over the same sorted diracs.  They are cut first and skipped by the
        // traversal: every dirac's plane bounds the cell, so the cut order changes, not the
        // cell.
        const bool use_hints = WARM && a.hint_col != nullptr;
        int hj = -1;
        if (use_hints && lane < HCAP3) {
            const int j = a.hint_col[HCAP3 * t + lane];
            if (j >= 0 && (int64_t)j < a.N && (int64_t)j != t) hj = j;
        }
        // After the warm start the polyhedron is nearly final, and with spread weights it
        // often lies away from its dirac: a candidate is then also tested against a sphere
        // about the vertices' mean, which holds the polyhedron as tightly as its own shape
        // allows.  Centre and squared radius stay in the face-area slots of the LDS slice
        // (free until the integration), so the cold passes keep their registers.
        bool centred = false;
        double *const sph = L.acc + (WV - 4);
        auto set_sphere = [&]() {
            double mx, my, mz, r2;
            C.sphere(mx, my, mz, r2);
            if (lane == 0) {
                sph[0] = mx;
                sph[1] = my;
                sph[2] = mz;
                sph[3] = r2;
            }
            wave_sync();
        };
        // true when the plane v.q <= h cannot cut the polyhedron (no vertex beyond it)
        auto cannot_cut = [&](double vx, double vy, double vz, double h, double vv) {
            if (h > 0 && R2 * vv <= h * h * (1 - c3::Eps3<T>::cand)) return true;   // sphere about the dirac
            if (!WARM || !centred) return false;
            const double d = h - (vx * sph[0] + vy * sph[1] + vz * sph[2]);
            return d > 0 && sph[3] * vv <= d * d * (1 - c3::Eps3<T>::cand);
        };
        // the lanes' candidates j (-1: none): tested against the current spheres, and only
        // the ballot of possible cutters applied, in lane order
        auto cut_lanes = [&](int j, bool skip_hints) {
            // every lane loads (a lane without a candidate loads the cell's own dirac): no
            // branch around the loads; the tests below are evaluated without short-circuits
            const int64_t jl = j >= 0 ? (int64_t)j : t;
            const double vx = a.spos[3 * jl] - c0x, vy = a.spos[3 * jl + 1] - c0y, vz = a.spos[3 * jl + 2] - c0z;
            const double vv = vx * vx + vy * vy + vz * vz;
            const double hv = 0.5 * (vv + w0 - a.sw[jl]);
            bool cand = (j >= 0) & (jl != t);
            if (cand & (vv == 0)) bad = true;
            cand = cand & (vv != 0);
            if (cand && cannot_cut(vx, vy, vz, hv, vv)) cand = false;
            unsigned long long mask = __ballot(cand);
            while (mask && C.nv > 0) {
                const int k = __ffsll((long long)mask) - 1;
                mask &= mask - 1;
                const double ux = lane_value(vx, k), uy = lane_value(vy, k), uz = lane_value(vz, k),
                             uh = lane_value(hv, k), uv = lane_value(vv, k);
                const int uj = lane_value(j, k);
                if (WARM && skip_hints && use_hints && __ballot(hj == uj)) continue;   // cut by the warm start
                if (cannot_cut(ux, uy, uz, uh, uv)) continue;                   // the polyhedron shrank
                const double tie = c3::Eps3<T>::tie * sqrt(uv * R2);
                const int rc = C.cut(T(ux), T(uy), T(uz), T(uh), uj, ux, uy, uz, uh, tie, true);
                if (rc == 2) {
                    overflow = true;
                    break;
                }
                if (rc == 1 && C.nv > 0) {
                    R2 = C.radius2();
                    if (WARM && centred) set_sphere();
                }
            }
        };
        if (WARM && use_hints && C.nv > 0 && !overflow) {
            cut_lanes(hj, false);
            if (C.nv > 0 && !overflow) {
                set_sphere();
                centred = true;
            }
        }
        // The diracs of the grid cells held by the lanes (ok, [cs, ce)) are staged 64 at a
        // time (one per lane, loaded in parallel) and handed to cut_lanes.
        auto process_cells = [&](bool ok, uint32_t cs, uint32_t ce) {
            const int cnt = ok ? (int)(ce - cs) : 0;
            const int pre = wave_excl_scan(cnt, lane);
            const int tot = __shfl(pre + cnt, WV - 1);
            if (tot == 0) return;
            L.cpre[lane] = pre;
            L.ccs[lane] = (int)cs;
            L.ccnt[lane] = cnt;
            wave_sync();
            for (int b0 = 0; b0 < tot && C.nv > 0 && !overflow; b0 += WV) {
                const int gi = b0 + lane;
                int j = -1;
                if (gi < tot) {
                    int lo = 0, hi = WV - 1;   // last cell slot with cpre <= gi and a non-empty range
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (L.cpre[mid] <= gi) lo = mid;
                        else hi = mid - 1;
                    }
                    while (L.ccnt[lo] == 0) --lo;
                    j = L.ccs[lo] + (gi - L.cpre[lo]);
                }
                cut_lanes(j, true);
            }
            wave_sync();
        };
        // Rings r0..r1 of grid cells around the dirac's cell, 64 cells at a time (one per
        // lane, tested in parallel).  `finished`: no dirac beyond the last ring can cut.
        bool finished = false;
        auto rings = [&](int r0, int r1) {
            for (int r = r0; r <= r1 && C.nv > 0 && !overflow; ++r) {
                if (r > 0) {
                    double dm = DBL_MAX;
                    bool any = false;
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        if (hh[d] - r + 1 > 0) { dm = fmin(dm, c0[d] - (g.lo[d] + (hh[d] - r + 1) * g.h[d])); any = true; }
                        if (hh[d] + r < g.ncell[d]) { dm = fmin(dm, (g.lo[d] + (hh[d] + r) * g.h[d]) - c0[d]); any = true; }
                    }
                    if (!any || c3::region_skippable3<T>(fmax(dm, 0.0), R2, w0, W)) {
                        finished = true;
                        return;
                    }
                }
                const int side = 2 * r + 1;
                const int ncube = side * side * side;
                for (int base = 0; base < ncube && C.nv > 0 && !overflow; base += WV) {
                    const int c = base + lane;
                    int dz = c / (side * side) - r, dy = (c / side) % side - r, dx = c % side - r;
                    bool ok = c < ncube && (dx == r || dx == -r || dy == r || dy == -r || dz == r || dz == -r);
                    const int gx = hh[0] + dx, gy = hh[1] + dy, gz = hh[2] + dz;
                    ok = ok && gx >= 0 && gy >= 0 && gz >= 0 && gx < g.ncell[0] && gy < g.ncell[1] && gz < g.ncell[2];
                    uint32_t cs = 0, ce = 0;
                    if (ok) {
                        const uint32_t code = morton3(gx, gy, gz);
                        cs = a.cell_start[code];
                        ce = a.cell_start[code + 1];
                        if (cs == ce) {
                            ok = false;
                        } else {
                            const double bx = g.lo[0] + gx * g.h[0], by = g.lo[1] + gy * g.h[1], bz = g.lo[2] + gz * g.h[2];
                            const double ex = fmax(0.0, fmax(bx - c0x, c0x - (bx + g.h[0])));
                            const double ey = fmax(0.0, fmax(by - c0y, c0y - (by + g.h[1])));
                            const double ez = fmax(0.0, fmax(bz - c0z, c0z - (bz + g.h[2])));
                            ok = !c3::region_skippable3<T>(sqrt(ex * ex + ey * ey + ez * ez), R2, w0, a.cell_wmax[code]);
                        }
                    }
                    process_cells(ok, cs, ce);
                }
            }
        };
        // Far field: the cells beyond the near rings, found by a breadth-first descent of
        // the max-weight pyramid (Morton nodes of 2^k cells per axis); a node is dropped
        // when no dirac under it can cut the current cell (its max weight, its box) or when
        // it lies inside the near rings.  Returns false when the frontier outgrows its LDS
        // capacity (nothing has been cut yet by this pass then).
        auto far_field = [&]() -> bool {
            int cur = 0, na = 1;
            if (lane == 0) L.fr[0][0] = 0u;
            wave_sync();
            for (int k = g.bits; k >= 1; --k) {
                const int span = 1 << (k - 1);
                const double *wl = a.wpyr + pyr_off(g.ncodes, 3, k - 1);
                int nb = 0;
                for (int base = 0; base < na * 8; base += WV) {
                    const int idx = base + lane;
                    bool keep = false;
                    uint32_t child = 0;
                    if (idx < na * 8) {
                        child = L.fr[cur][idx >> 3] * 8u + (uint32_t)(idx & 7);
                        const int cc[3] = {(int)compact3(child) * span, (int)compact3(child >> 1) * span,
                                           (int)compact3(child >> 2) * span};
                        if (cc[0] < g.ncell[0] && cc[1] < g.ncell[1] && cc[2] < g.ncell[2]) {
                            bool inside = true;
#pragma unroll
                            for (int d = 0; d < 3; ++d)
                                inside = inside && cc[d] >= hh[d] - NEAR_RINGS && cc[d] + span - 1 <= hh[d] + NEAR_RINGS;
                            const double wm = wl[child];
                            if (!inside && wm > -DBL_MAX) {
                                double e2 = 0;
#pragma unroll
                                for (int d = 0; d < 3; ++d) {
                                    const int c1 = cc[d] + span < g.ncell[d] ? cc[d] + span : g.ncell[d];
                                    const double lo = g.lo[d] + cc[d] * g.h[d], hi = g.lo[d] + c1 * g.h[d];
                                    const double e = fmax(0.0, fmax(lo - c0[d], c0[d] - hi));
                                    e2 += e * e;
                                }
                                keep = !c3::region_skippable3<T>(sqrt(e2), R2, w0, wm);
                            }
                        }
                    }
                    const unsigned long long m = __ballot(keep);
                    const int pos = nb + __popcll(m & ((1ull << lane) - 1ull));
                    if (keep && pos < FRONT) L.fr[cur ^ 1][pos] = child;
                    nb += __popcll(m);
                }
                if (nb > FRONT) return false;
                wave_sync();
                cur ^= 1;
                na = nb;
                if (na == 0) return true;
            }
            // level 0: grid cells
            for (int base = 0; base < na && C.nv > 0 && !overflow; base += WV) {
                const int idx = base + lane;
                uint32_t cs = 0, ce = 0;
                bool ok = idx < na;
                if (ok) {
                    const uint32_t code = L.fr[cur][idx];
                    cs = a.cell_start[code];
                    ce = a.cell_start[code + 1];
                    ok = cs != ce;
                }
                process_cells(ok, cs, ce);
            }
            return true;
        };
        const bool pyramid = a.wpyr != nullptr && g.bits >= 2;
        rings(0, pyramid ? NEAR_RINGS : maxr);
        if (pyramid && !finished && C.nv > 0 && !overflow && !far_field()) rings(NEAR_RINGS + 1, maxr);
        if (C.nv > 0 && !overflow) C.compact_planes();   // voxel cuts then start from the face count
        if (bad && lane == 0) flag_or(a.flags, FLAG_BAD_INPUT);
        if (overflow) {
            if (lane == 0) {
                unsigned k = atomicAdd(a.ovf_count, 1u);
                a.ovf_list[k] = (int)t;
            }
            wave_sync();
            continue;
        }
        // ---- integrals ----
        double mass = 0, mom[3] = {0, 0, 0}, off[3];
        double *const want = a.mom ? off : nullptr;
        const int npc = C.np;
        if (lane < npc) L.hacc[lane] = 0;
        wave_sync();
        if (C.nv > 0) {
            if (!a.img && a.nitems < 2) {
                const double vol = C.integrate(want);
                mass = a.coeff * vol;
                if (want)
                    for (int d = 0; d < 3; ++d) mom[d] = mass * off[d];
                if (lane < npc) L.hacc[lane] = a.coeff * L.acc[lane];
            } else if (!a.img) {
                // several domain items (for_each_intersection, ConvexPolyhedronAssembly.tcc:132-142):
                // the cell cut by each item's box sides and planes, integrated with its density
                double X, Y, Z;
                C.vertex64(X, Y, Z);
                const bool valid = lane < C.nv;
                double mn[3], mx[3];
                mn[0] = -wave_max_dpp(valid ? -X : -DBL_MAX); mx[0] = wave_max_dpp(valid ? X : -DBL_MAX);
                mn[1] = -wave_max_dpp(valid ? -Y : -DBL_MAX); mx[1] = wave_max_dpp(valid ? Y : -DBL_MAX);
                mn[2] = -wave_max_dpp(valid ? -Z : -DBL_MAX); mx[2] = wave_max_dpp(valid ? Z : -DBL_MAX);
                const int nv0 = C.nv, np0 = C.np;
                const T sx = C.x, sy = C.y, sz = C.z;
                const int stri = C.tri;
                for (int k = 0; k < a.nitems && !overflow; ++k) {
                    const double *b = a.it_box + 6 * k;
                    C.vlo0 = b[0] - c0[0];
                    C.vlo1 = b[1] - c0[1];
                    C.vlo2 = b[2] - c0[2];
                    C.vhi0 = b[3] - c0[0];
                    C.vhi1 = b[4] - c0[1];
                    C.vhi2 = b[5] - c0[2];
                    if (C.vlo0 > mx[0] || C.vhi0 < mn[0] || C.vlo1 > mx[1] || C.vhi1 < mn[1] || C.vlo2 > mx[2] ||
                        C.vhi2 < mn[2])
                        continue;   // the item's box misses the cell
                    bool empty = false;
#pragma unroll
                    for (int sd = 0; sd < 6; ++sd) {
                        const int axis = sd >> 1, up = sd & 1;
                        if (empty || overflow) continue;
                        const double bound = up ? C.vhi(axis) : C.vlo(axis);
                        if (up ? !(bound < mx[axis]) : !(bound > mn[axis])) continue;   // cannot cut
                        const double nxx = axis == 0 ? (up ? 1.0 : -1.0) : 0.0, nyy = axis == 1 ? (up ? 1.0 : -1.0) : 0.0,
                                     nzz = axis == 2 ? (up ? 1.0 : -1.0) : 0.0, hq = up ? bound : -bound;
                        const int rc = C.cut(T(nxx), T(nyy), T(nzz), T(hq), -7 - sd, nxx, nyy, nzz, hq,
                                             c3::Eps3<T>::tie * sqrt(R2), false);
                        if (rc == 2) { overflow = true; break; }
                        if (C.nv == 0) empty = true;
                    }
                    for (int q = a.it_off[k]; q < a.it_off[k + 1] && !empty && !overflow; ++q) {
                        const double *pl = a.it_planes + 4 * q;
                        const double hl = pl[3] - (pl[0] * c0[0] + pl[1] * c0[1] + pl[2] * c0[2]);
                        const double nn = sqrt(pl[0] * pl[0] + pl[1] * pl[1] + pl[2] * pl[2]);
                        const int rc = C.cut(T(pl[0]), T(pl[1]), T(pl[2]), T(hl), -100 - q, pl[0], pl[1], pl[2], hl,
                                             c3::Eps3<T>::tie * nn * sqrt(R2), false);
                        if (rc == 2) { overflow = true; break; }
                        if (C.nv == 0) empty = true;
                    }
                    if (!overflow && !empty) {
                        const double rho = a.it_coeff[k];
                        const double vol = C.integrate(want);
                        mass += rho * vol;
                        if (want)
                            for (int d = 0; d < 3; ++d) mom[d] += rho * vol * off[d];
                        if (lane < npc) L.hacc[lane] += rho * L.acc[lane];
                    }
                    // restore the cell
                    C.nv = nv0;
                    C.np = np0;
                    C.x = sx;
                    C.y = sy;
                    C.z = sz;
                    C.tri = stri;
                    C.write_edges();
                }
                if (overflow) {
                    if (lane == 0) {
                        unsigned kq = atomicAdd(a.ovf_count, 1u);
                        a.ovf_list[kq] = (int)t;
                    }
                    wave_sync();
                    continue;
                }
            } else {
                double X, Y, Z;
                C.vertex64(X, Y, Z);
                const bool valid = lane < C.nv;
                double mn[3], mx[3];
                mn[0] = -wave_max(valid ? -X : -DBL_MAX); mx[0] = wave_max(valid ? X : -DBL_MAX);
                mn[1] = -wave_max(valid ? -Y : -DBL_MAX); mx[1] = wave_max(valid ? Y : -DBL_MAX);
                mn[2] = -wave_max(valid ? -Z : -DBL_MAX); mx[2] = wave_max(valid ? Z : -DBL_MAX);
                int i0[3], i1[3];
                double vs[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    vs[d] = (a.dom_hi[d] - a.dom_lo[d]) / a.img_n[d];
                    const double lo = mn[d] + c0[d] - a.dom_lo[d], hi = mx[d] + c0[d] - a.dom_lo[d];
                    const int u0 = (int)floor(lo / vs[d]), u1 = (int)floor(hi / vs[d]);
                    i0[d] = u0 < 0 ? 0 : (u0 >= a.img_n[d] ? a.img_n[d] - 1 : u0);
                    i1[d] = u1 < 0 ? 0 : (u1 >= a.img_n[d] ? a.img_n[d] - 1 : u1);
                }
                const int nv0 = C.nv, np0 = C.np;
                const T sx = C.x, sy = C.y, sz = C.z;
                const int stri = C.tri;
                for (int iz = i0[2]; iz <= i1[2] && !overflow; ++iz)
                    for (int iy = i0[1]; iy <= i1[1] && !overflow; ++iy)
                        for (int ix = i0[0]; ix <= i1[0] && !overflow; ++ix) {
                            const int iv[3] = {ix, iy, iz};
                            const double rho = a.img[((int64_t)iz * a.img_n[1] + iy) * a.img_n[0] + ix];
                            C.vlo0 = a.dom_lo[0] + vs[0] * ix - c0[0];
                            C.vhi0 = a.dom_lo[0] + vs[0] * (ix + 1) - c0[0];
                            C.vlo1 = a.dom_lo[1] + vs[1] * iy - c0[1];
                            C.vhi1 = a.dom_lo[1] + vs[1] * (iy + 1) - c0[1];
                            C.vlo2 = a.dom_lo[2] + vs[2] * iz - c0[2];
                            C.vhi2 = a.dom_lo[2] + vs[2] * (iz + 1) - c0[2];
                            bool empty = false;
#pragma unroll
                            for (int s = 0; s < 6; ++s) {
                                const int axis = s >> 1, up = s & 1;
                                if (empty || overflow) continue;
                                if (up ? (iv[axis] == i1[axis]) : (iv[axis] == i0[axis])) continue;
                                const double off = up ? C.vhi(axis) : C.vlo(axis);
                                const double nxx = axis == 0 ? (up ? 1.0 : -1.0) : 0.0, nyy = axis == 1 ? (up ? 1.0 : -1.0) : 0.0,
                                             nzz = axis == 2 ? (up ? 1.0 : -1.0) : 0.0, hq = up ? off : -off;
                                const int rc = C.cut(T(nxx), T(nyy), T(nzz), T(hq), -7 - s, nxx, nyy, nzz, hq,
                                                     c3::Eps3<T>::tie * sqrt(R2), false);
                                if (rc == 2) { overflow = true; break; }
                                if (C.nv == 0) empty = true;
                            }
                            if (!overflow && !empty) {
                                const double vol = C.integrate(want);
                                mass += rho * vol;
                                if (want)
                                    for (int d = 0; d < 3; ++d) mom[d] += rho * vol * off[d];
                                if (lane < npc) L.hacc[lane] += rho * L.acc[lane];
                            }
                            // restore the cell
                            C.nv = nv0;
                            C.np = np0;
                            C.x = sx;
                            C.y = sy;
                            C.z = sz;
                            C.tri = stri;
                            C.write_edges();
                        }
                if (overflow) {
                    if (lane == 0) {
                        unsigned k = atomicAdd(a.ovf_count, 1u);
                        a.ovf_list[k] = (int)t;
                    }
                    wave_sync();
                    continue;
                }
            }
        }
        wave_sync();
        // ---- Hessian row, boundary measure ----
        const unsigned long long used = C.nv > 0 ? C.used_planes() : 0ull;
        const bool mine = lane < npc && ((used >> lane) & 1ull);
        const int id = lane < npc ? L.pid[lane] : -1;
        const double ar = mine ? L.hacc[lane] : 0.0;
        const bool nb = mine && id >= 0;
        double b = 0;
        if (nb) {
            double nx, ny, nz, hv;
            C.plane64(lane, nx, ny, nz, hv);
            b = 0.5 * ar / sqrt(nx * nx + ny * ny + nz * nz);
        }
        const int slot = wave_excl_scan_small(nb ? 1 : 0);
        const int nrow = lane_value(slot + (nb ? 1 : 0), WV - 1);
        if (nb && slot < a.ell_cap) {
            a.ell_col[slot * a.ell_stride + t] = id;
            a.ell_val[slot * a.ell_stride + t] = -b;
        }
        const double der = wave_sum_dpp(b);
        const double perim = wave_sum_dpp(mine && ((!a.img && a.nitems < 2) || id >= 0) ? ar : 0.0);
        if (lane == 0) {
            a.mass[t] = mass;
            if (a.mom)
                for (int d = 0; d < 3; ++d) a.mom[3 * t + d] = mom[d];
            a.arc[t] = 0;
            a.perim[t] = perim;
            a.diag[t] = der;
            a.row_len[t] = nrow < a.ell_cap ? nrow : a.ell_cap;
            if (nrow > a.ell_cap) flag_or(a.flags, FLAG_ROW_OVERFLOW);
            if (mass == 0) flag_or(a.flags, FLAG_EMPTY);
            if (a.ctr) {
                const unsigned long long K = __popcll(used);
                const int sl = (int)(t % WCTR);
                const unsigned long long F = 12 * K * K + 100 * K;
                ctr_add(a.ctr + sl, F > 284 ? F - 284 : 0);
                ctr_add(a.ctr + WCTR + sl, K);
            }
        }
        wave_sync();
    }
}

// hint_col[HCAP3 t + e] = sorted neighbour e of row t (-1 past the row), one thread per entry
__global__ __launch_bounds__(256) void k_hint_copy(const int *row_len, const int *ell_col, int64_t stride, int cap,
                                                  int64_t n, int *hint_col) {
    const int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x;
    if (i >= n * HCAP3) return;
    const int64_t t = i / HCAP3;
    const int e = (int)(i % HCAP3);
    int L = row_len[t];
    L = L < cap ? L : cap;
    hint_col[i] = e < L ? ell_col[e * stride + t] : -1;
}

template <class T>
void launch_wave3(sdot_pd *pd, const CellArgs &a) {
    int dev = 0, ncu = 0;
    HIPCK(hipGetDevice(&dev));
    HIPCK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    int64_t nblocks = (int64_t)ncu * 8;
    const int64_t need = (a.N + WPB - 1) / WPB;
    if (nblocks > need) nblocks = need;
    if (nblocks < 1) return;
    const CellArgs *da = upload_args(pd, a, 0);
    if (a.hint_col)
        hipLaunchKernelGGL((k_cells3_wave<T, true>), dim3((unsigned)nblocks), dim3(WV * WPB), 0, pd->stream, da);
    else
        hipLaunchKernelGGL((k_cells3_wave<T, false>), dim3((unsigned)nblocks), dim3(WV * WPB), 0, pd->stream, da);
    HIPCK(hipGetLastError());
}

}  // namespace

void snapshot_hints(sdot_pd *pd) {
    const int64_t N = pd->N;
    pd->hint_col.ensure((size_t)HCAP3 * N);
    const int64_t tot = (int64_t)HCAP3 * N;
    hipLaunchKernelGGL(k_hint_copy, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, pd->stream, pd->row_len.p,
                       pd->ell_col.p, (int64_t)N, pd->ell_cap, N, pd->hint_col.p);
    HIPCK(hipGetLastError());
}

void run_cells3(sdot_pd *pd, CellArgs &a) {
    hipStream_t s = pd->stream;
    a.work = nullptr;
    a.work_n = 0;
    const bool mixed = pd->precision == SDOT_MIXED;
    // position derivatives: the thread kernel in FP64, with per-face moments
    if (a.jdiag) {
        launch_thread3<double, VCAP3, PCAP3, false, true>(pd, a);
        HIPCK(hipEventRecord(pd->ev[3], s));
        unsigned c1 = 0;
        HIPCK(hipMemcpyAsync(&c1, pd->flags.p + 1, sizeof(unsigned), hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        pd->last.nb_cells_overflow = c1;
        if (c1 > 0) {
            int *list2 = pd->ovf_list.p + a.N + 1;
            HIPCK(hipMemcpyAsync(list2, pd->ovf_list.p, c1 * sizeof(int), hipMemcpyDeviceToDevice, s));
            a.work = list2;
            a.work_n = c1;
            launch_thread3<double, VCAP3W, PCAP3W, true, true>(pd, a);
            a.work = nullptr;
        }
        return;
    }
    // translated diracs (periodic images) or the R2 integrals: every cell on the thread
    // kernel, whose traversal visits the grid once per translation; otherwise the
    // wavefront kernel
    if (a.ntrans > 0 || a.rf) {
        if (mixed) launch_thread3<float, VCAP3, PCAP3, false>(pd, a);
        else launch_thread3<double, VCAP3, PCAP3, false>(pd, a);
    } else {
        a.hint_col = pd->warm && pd->hint_ok && pd->hint_col.n >= (size_t)HCAP3 * a.N ? pd->hint_col.p : nullptr;
        if (mixed) launch_wave3<float>(pd, a);
        else launch_wave3<double>(pd, a);
        a.hint_col = nullptr;
    }
    HIPCK(hipEventRecord(pd->ev[3], s));
    unsigned cnt[2] = {0, 0};
    HIPCK(hipMemcpyAsync(cnt, pd->flags.p, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, s));
    HIPCK(hipStreamSynchronize(s));
    pd->last.nb_cells_overflow = cnt[1];
    // cells beyond 64 vertices / planes: the thread kernel, tier 1 then tier 2; each
    // tier works on a copy of the list while the next tier's list refills ovf_list
    int *list2 = pd->ovf_list.p + a.N + 1;
    for (int tier = a.ntrans > 0 || a.rf ? 2 : 1; tier <= 2 && cnt[1] > 0; ++tier) {
        HIPCK(hipMemcpyAsync(list2, pd->ovf_list.p, cnt[1] * sizeof(int), hipMemcpyDeviceToDevice, s));
        fill_u32(pd->flags.p + 1, 1, 0u, s);
        a.work = list2;
        a.work_n = cnt[1];
        if (tier == 1) {
            if (mixed) launch_thread3<float, VCAP3, PCAP3, false>(pd, a);
            else launch_thread3<double, VCAP3, PCAP3, false>(pd, a);
        } else {
            if (mixed) launch_thread3<float, VCAP3W, PCAP3W, true>(pd, a);
            else launch_thread3<double, VCAP3W, PCAP3W, true>(pd, a);
        }
        HIPCK(hipMemcpyAsync(&cnt[1], pd->flags.p + 1, sizeof(unsigned), hipMemcpyDeviceToHost, s));
        HIPCK(hipStreamSynchronize(s));
        a.work = nullptr;
        if (tier == 1) pd->last.nb_cells_overflow += cnt[1];
    }
}

}  // namespace sdot_amd
