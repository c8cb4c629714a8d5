// Plain FP64 tile kernels for gfx950: the SSP-RK3 step of TracerAdvection.advect with x, y and z WENO5 sweeps, and
// the standalone helpers (WENO faces, upwind faces, Zalesak coefficients, open-edge fluxes, periodic halo, content
// sums), on the masked C-grid (reference: core/dynamics/sigma_baroclinic.py:265-449, 856-875;
// core/components/tracers.py:65-193, 342-375).  The FCT step itself runs in the wave-marching kernels.
//
// Layout (C-contiguous, k = 0 at the bed): stacks (nt, n, ny, nx); cells V0, V1, ok (n, ny, nx); faces
// Qx (n, ny, nx+1), Qy (n, ny+1, nx), Fz (n+1, ny, nx); face stacks fx (nt, n, ny, nx+1), fy (nt, n, ny+1, nx).
//
// Every block owns a TX x TY tile of one (tracer, layer) plane.  It stages the input tile with a 3-cell halo in
// LDS (x-contiguous rows, one 64-lane wavefront per row for TX = 64), reconstructs each face of the tile once into
// LDS, and differences the faces of its own cells.  A face on a tile border is computed by both blocks that touch
// it (for the divergence) but written to global memory only by its owner (faces lo..lo+TX-1 of the tile, plus the
// domain's last face), so read-modify-write of face stacks never races.
#pragma once
#include "ocx_weno.hpp"

namespace ocx {

// ------------------------------------------------------------------------------------------------ tile loads
template <int TX, int TY, int H>
__device__ __forceinline__ void load_tile(double* sc, uint8_t* so, const double* __restrict__ src,
                                          const uint8_t* __restrict__ okp, int i0, int j0, int ny, int nx) {
    constexpr int LX = TX + 2 * H, LY = TY + 2 * H;
    const int tid = threadIdx.y * TX + threadIdx.x;
    for (int idx = tid; idx < LY * LX; idx += TX * TY) {
        const int ly = idx / LX, lx = idx - ly * LX;
        const int j = j0 - H + ly, i = i0 - H + lx;
        const bool in = j >= 0 && j < ny && i >= 0 && i < nx;
        sc[idx] = in ? src[(size_t)j * nx + i] : 0.0;
        if (so) so[idx] = in ? okp[(size_t)j * nx + i] : (uint8_t)0;
    }
}

// Face fluxes q * c_face of the tile's faces into LDS (sfx: TY rows x TX+1 faces; sfy: TY+1 faces x TX columns),
// from the input tile sc/so (halo 3).  WENO (SCH_JS, SCH_Z), rules of _advect_faces_h (sigma_baroclinic.py:
// 265-321): interior faces with q == 0 stay 0; open edges take ext on inflow, the edge cell on outflow; closed
// edges 0.  UP3 / UP1 / centred (SCH_UP3, SCH_UP1, SCH_C6), rules of rhs_numpy (tracers.py:139-151): every
// interior face is q times the _up3_faces or _centred_faces value, a closed edge is q * 0, open edges as WENO.
// ext == nullptr: the input is the exterior.
// Face items per thread of a TX x TY tile: x faces TY rows x TX+1, y faces TY+1 rows x TX, item m of thread tid
// at index tid + m * TX * TY.
template <int TX, int TY>
struct TileItems {
    static constexpr int NT = TX * TY, NX = (TY * (TX + 1) + NT - 1) / NT, NY = ((TY + 1) * TX + NT - 1) / NT;
};

// The face transports of the thread's items (0 outside the domain), for a caller that loads them ahead.
template <int TX, int TY>
__device__ __forceinline__ void tile_face_q(double* qx, double* qy, const double* __restrict__ Qxk,
                                            const double* __restrict__ Qyk, int i0, int j0, int ny, int nx) {
    using IT = TileItems<TX, TY>;
    const int tid = threadIdx.y * TX + threadIdx.x;
#pragma unroll
    for (int m = 0; m < IT::NX; ++m) {
        const int idx = tid + m * IT::NT, ly = idx / (TX + 1), lf = idx - ly * (TX + 1);
        const int j = j0 + ly, f = i0 + lf;
        qx[m] = (idx < TY * (TX + 1) && j < ny && f <= nx) ? Qxk[(size_t)j * (nx + 1) + f] : 0.0;
    }
#pragma unroll
    for (int m = 0; m < IT::NY; ++m) {
        const int idx = tid + m * IT::NT, lf = idx / TX, lx = idx - lf * TX;
        const int f = j0 + lf, i = i0 + lx;
        qy[m] = (idx < (TY + 1) * TX && i < nx && f <= ny) ? Qyk[(size_t)f * nx + i] : 0.0;
    }
}

// qx / qy (optional): the transports of the thread's items from tile_face_q instead of loads from Qxk / Qyk.
template <int SCH, int TX, int TY>
__device__ __forceinline__ void tile_faces(double* sfx, double* sfy, const double* sc, const uint8_t* so,
                                                const double* __restrict__ Qxk, const double* __restrict__ Qyk,
                                                const double* __restrict__ extp, int i0, int j0, int ny, int nx,
                                                unsigned flags, const double* qx = nullptr,
                                                const double* qy = nullptr) {
    constexpr int H = 3, LX = TX + 2 * H;
    using IT = TileItems<TX, TY>;
    const int tid = threadIdx.y * TX + threadIdx.x;
#pragma unroll
    for (int m = 0; m < IT::NX; ++m) {
        const int idx = tid + m * IT::NT;
        if (idx >= TY * (TX + 1)) break;
        const int ly = idx / (TX + 1), lf = idx - ly * (TX + 1);
        const int j = j0 + ly, f = i0 + lf;
        double val = 0.0;
        if (j < ny && f <= nx) {
            const double q = qx ? qx[m] : Qxk[(size_t)j * (nx + 1) + f];
            const double* row = sc + (ly + H) * LX;
            const uint8_t* orow = so + (ly + H) * LX;
            if (f > 0 && f < nx) {
                if constexpr (SCH == SCH_C6)
                    val = q * centred_face(row, orow, 1, lf + H);
                else if constexpr (SCH >= SCH_UP3)
                    val = q * up_face<SCH == SCH_UP3>(row, orow, 1, lf + H, q);
                else if (q != 0.0)
                    val = q * lds_face<SCH == SCH_Z>(row, orow, 1, lf + H, q);
            } else if (f == 0) {
                if (flags & OPEN_W) {
                    const double cc = row[lf + H];
                    const double e = extp ? extp[(size_t)j * nx] : cc;
                    val = q * (q > 0.0 ? e : cc);
                } else if constexpr (SCH >= SCH_UP3) {
                    val = q * 0.0;
                }
            } else if (flags & OPEN_E) {
                const double cc = row[lf + H - 1];
                const double e = extp ? extp[(size_t)j * nx + nx - 1] : cc;
                val = q * (q < 0.0 ? e : cc);
            } else if constexpr (SCH >= SCH_UP3) {
                val = q * 0.0;
            }
        }
        sfx[idx] = val;
    }
#pragma unroll
    for (int m = 0; m < IT::NY; ++m) {
        const int idx = tid + m * IT::NT;
        if (idx >= (TY + 1) * TX) break;
        const int lf = idx / TX, lx = idx - lf * TX;
        const int f = j0 + lf, i = i0 + lx;
        double val = 0.0;
        if (i < nx && f <= ny) {
            const double q = qy ? qy[m] : Qyk[(size_t)f * nx + i];
            const double* col = sc + lx + H;
            const uint8_t* ocol = so + lx + H;
            if (f > 0 && f < ny) {
                if constexpr (SCH == SCH_C6)
                    val = q * centred_face(col, ocol, LX, lf + H);
                else if constexpr (SCH >= SCH_UP3)
                    val = q * up_face<SCH == SCH_UP3>(col, ocol, LX, lf + H, q);
                else if (q != 0.0)
                    val = q * lds_face<SCH == SCH_Z>(col, ocol, LX, lf + H, q);
            } else if (f == 0) {
                if (flags & OPEN_S) {
                    const double cc = col[(lf + H) * LX];
                    const double e = extp ? extp[i] : cc;
                    val = q * (q > 0.0 ? e : cc);
                } else if constexpr (SCH >= SCH_UP3) {
                    val = q * 0.0;
                }
            } else if (flags & OPEN_N) {
                const double cc = col[(lf + H - 1) * LX];
                const double e = extp ? extp[(size_t)(ny - 1) * nx + i] : cc;
                val = q * (q < 0.0 ? e : cc);
            } else if constexpr (SCH >= SCH_UP3) {
                val = q * 0.0;
            }
        }
        sfy[idx] = val;
    }
}

// One x-face (tile row ly, face lf = 0 .. TX of the tile's row, the face between columns i0+lf-1 and i0+lf) or one
// y-face (face row lf = 0 .. TY, column lx) of the LDS tile, with the same rules as tile_faces.  INNER: the tile's
// faces are all interior faces of the domain (no edge rules, no bounds tests).
template <int SCH, int TX, int TY, bool INNER = false>
__device__ __forceinline__ double tile_xface(const double* sc, const uint8_t* so, const double* __restrict__ extp,
                                             int ly, int lf, int i0, int j0, int ny, int nx, unsigned flags, double q,
                                             bool known) {
    constexpr int H = 3, LX = TX + 2 * H;
    const int j = j0 + ly, f = i0 + lf;
    double val = 0.0;
    if (INNER || (j < ny && f <= nx)) {
        const double* row = sc + (ly + H) * LX;
        const uint8_t* orow = so + (ly + H) * LX;
        if (INNER || (f > 0 && f < nx)) {
            if constexpr (SCH == SCH_C6)
                val = q * centred_face(row, orow, 1, lf + H);
            else if constexpr (SCH >= SCH_UP3)
                val = q * up_face<SCH == SCH_UP3>(row, orow, 1, lf + H, q);
            else if (q != 0.0)
                val = q * lds_face<SCH == SCH_Z>(row, orow, 1, lf + H, q, known);
        } else if (f == 0) {
            if (flags & OPEN_W) {
                const double cc = row[lf + H];
                const double e = extp ? extp[(size_t)j * nx] : cc;
                val = q * (q > 0.0 ? e : cc);
            } else if constexpr (SCH >= SCH_UP3) {
                val = q * 0.0;
            }
        } else if (flags & OPEN_E) {
            const double cc = row[lf + H - 1];
            const double e = extp ? extp[(size_t)j * nx + nx - 1] : cc;
            val = q * (q < 0.0 ? e : cc);
        } else if constexpr (SCH >= SCH_UP3) {
            val = q * 0.0;
        }
    }
    return val;
}

template <int SCH, int TX, int TY, bool INNER = false>
__device__ __forceinline__ double tile_yface(const double* sc, const uint8_t* so, const double* __restrict__ extp,
                                             int lf, int lx, int i0, int j0, int ny, int nx, unsigned flags, double q,
                                             bool known) {
    constexpr int H = 3, LX = TX + 2 * H;
    const int f = j0 + lf, i = i0 + lx;
    double val = 0.0;
    if (INNER || (i < nx && f <= ny)) {
        const double* col = sc + lx + H;
        const uint8_t* ocol = so + lx + H;
        if (INNER || (f > 0 && f < ny)) {
            if constexpr (SCH == SCH_C6)
                val = q * centred_face(col, ocol, LX, lf + H);
            else if constexpr (SCH >= SCH_UP3)
                val = q * up_face<SCH == SCH_UP3>(col, ocol, LX, lf + H, q);
            else if (q != 0.0)
                val = q * lds_face<SCH == SCH_Z>(col, ocol, LX, lf + H, q, known);
        } else if (f == 0) {
            if (flags & OPEN_S) {
                const double cc = col[(lf + H) * LX];
                const double e = extp ? extp[i] : cc;
                val = q * (q > 0.0 ? e : cc);
            } else if constexpr (SCH >= SCH_UP3) {
                val = q * 0.0;
            }
        } else if (flags & OPEN_N) {
            const double cc = col[(lf + H - 1) * LX];
            const double e = extp ? extp[(size_t)(ny - 1) * nx + i] : cc;
            val = q * (q < 0.0 ? e : cc);
        } else if constexpr (SCH >= SCH_UP3) {
            val = q * 0.0;
        }
    }
    return val;
}

// The faces of a narrow tile (TX x TY = 256 cells, TX + TY <= 64; blockDim (64, 4), thread t on cell (t / TX,
// t % TX)): each thread forms the west face and the north face of its cell; the tile's south faces (TX) and east
// faces (TY) are one item list for one wavefront, which rotates with the layer, so every wave forms the same number
// of faces over four layers and the tile's extra faces take one face evaluation instead of two.  In an interior
// tile the extra items differ only in their LDS line (a column at stride LX for a south face, a row at stride 1 for
// an east face), so they run as one code path; an edge tile forms them through tile_yface / tile_xface.
template <int SCH, int TX, int TY, bool INNER>
__device__ __forceinline__ void tile_faces_narrow(double* sfx, double* sfy, const double* sc, const uint8_t* so,
                                                  const double* __restrict__ extp, int i0, int j0, int ny, int nx,
                                                  unsigned flags, int k, const double* q4, bool known) {
    static_assert(TX * TY == 256 && TX + TY <= 64, "a 256-cell tile whose extra faces fit one wavefront");
    constexpr int H = 3, LX = TX + 2 * H;
    const int t = threadIdx.y * 64 + threadIdx.x, r = t / TX, c = t - r * TX;
    sfx[r * (TX + 1) + c] = tile_xface<SCH, TX, TY, INNER>(sc, so, extp, r, c, i0, j0, ny, nx, flags, q4[0], known);
    sfy[(r + 1) * TX + c] = tile_yface<SCH, TX, TY, INNER>(sc, so, extp, r + 1, c, i0, j0, ny, nx, flags, q4[1],
                                                           known);
    const int w = (int)__builtin_amdgcn_readfirstlane(threadIdx.y), lane = threadIdx.x;
    if (w != (k & 3) || lane >= TX + TY) return;
    const bool south = lane < TX;
    double v;
    if constexpr (INNER) {
        const int off = south ? lane + H : (lane - TX + H) * LX, st = south ? LX : 1, cc = south ? H : TX + H;
        const double q = q4[2];
        if constexpr (SCH == SCH_C6)
            v = q * centred_face(sc + off, so + off, st, cc);
        else if constexpr (SCH >= SCH_UP3)
            v = q * up_face<SCH == SCH_UP3>(sc + off, so + off, st, cc, q);
        else
            v = q != 0.0 ? q * lds_face<SCH == SCH_Z>(sc + off, so + off, st, cc, q, known) : 0.0;
    } else {
        v = south ? tile_yface<SCH, TX, TY>(sc, so, extp, 0, lane, i0, j0, ny, nx, flags, q4[2], known)
                  : tile_xface<SCH, TX, TY>(sc, so, extp, lane - TX, TX, i0, j0, ny, nx, flags, q4[2], known);
    }
    if (south)
        sfy[lane] = v;
    else
        sfx[(lane - TX) * (TX + 1) + TX] = v;
}

// The faces of the tile with one wavefront per tile row (TX = 64, blockDim (64, TY)): wave w forms the west faces
// of row w and the north faces of row w; the tile's last x-faces (east of each row) and its first y-faces (south of
// row 0) go to two waves that alternate with the layer parity, so every wave forms the same number of faces over
// two layers.  Narrower tiles go to tile_faces_narrow.  q*: the transports of those items, loaded ahead
// (tile_face_q_waves).  known: every value of the LDS tile is 0 or of magnitude in [2^-400, 2^100], so the WENO
// faces skip their per-face range test.
template <int SCH, int TX, int TY, bool INNER = false>
__device__ __forceinline__ void tile_faces_waves(double* sfx, double* sfy, const double* sc, const uint8_t* so,
                                                 const double* __restrict__ extp, int i0, int j0, int ny, int nx,
                                                 unsigned flags, int k, const double* q4, bool known) {
    if constexpr (TX < 64) {
        tile_faces_narrow<SCH, TX, TY, INNER>(sfx, sfy, sc, so, extp, i0, j0, ny, nx, flags, k, q4, known);
        return;
    }
    static_assert(TX < 64 || TY >= 2, "one wavefront per tile row");
    const int w = (int)__builtin_amdgcn_readfirstlane(threadIdx.y), lane = threadIdx.x;
    sfx[w * (TX + 1) + lane] =
        tile_xface<SCH, TX, TY, INNER>(sc, so, extp, w, lane, i0, j0, ny, nx, flags, q4[0], known);
    sfy[(w + 1) * TX + lane] =
        tile_yface<SCH, TX, TY, INNER>(sc, so, extp, w + 1, lane, i0, j0, ny, nx, flags, q4[1], known);
    const int ws = (k & 1) ? TY - 2 : 0, we = (k & 1) ? TY - 1 : 1;
    if (w == ws)
        sfy[lane] = tile_yface<SCH, TX, TY, INNER>(sc, so, extp, 0, lane, i0, j0, ny, nx, flags, q4[2], known);
    if (w == we && lane < TY)
        sfx[lane * (TX + 1) + TX] =
            tile_xface<SCH, TX, TY, INNER>(sc, so, extp, lane, TX, i0, j0, ny, nx, flags, q4[3], known);
}

// The transports of tile_faces_waves's items for one layer: Qx on the west face of the thread's cell, Qy on its north
// face, Qy on the south face of its column in row 0, Qx on the east face of tile row `lane` (lanes < TY); 0 outside
// the domain.
template <int TX, int TY>
__device__ __forceinline__ void tile_face_q_waves(double* q4, const double* __restrict__ Qxk,
                                                  const double* __restrict__ Qyk, int i0, int j0, int ny, int nx) {
    if constexpr (TX < 64) {
        // narrow tile: the thread's cell's west and north faces, and the item of the extra list (lane < TX: south
        // face of column lane; TX <= lane < TX + TY: east face of row lane - TX)
        const int t = threadIdx.y * 64 + threadIdx.x, r = t / TX, c = t - r * TX, lane = threadIdx.x;
        const int j = j0 + r, i = i0 + c;
        q4[0] = (j < ny && i <= nx) ? Qxk[(size_t)j * (nx + 1) + i] : 0.0;
        q4[1] = (i < nx && j + 1 <= ny) ? Qyk[(size_t)(j + 1) * nx + i] : 0.0;
        if (lane < TX)
            q4[2] = (i0 + lane < nx && j0 <= ny) ? Qyk[(size_t)j0 * nx + i0 + lane] : 0.0;
        else if (lane < TX + TY)
            q4[2] = (j0 + lane - TX < ny && i0 + TX <= nx) ? Qxk[(size_t)(j0 + lane - TX) * (nx + 1) + i0 + TX] : 0.0;
        else
            q4[2] = 0.0;
        q4[3] = 0.0;
        return;
    }
    const int w = threadIdx.y, lane = threadIdx.x;
    const int j = j0 + w, i = i0 + lane;
    q4[0] = (j < ny && i <= nx) ? Qxk[(size_t)j * (nx + 1) + i] : 0.0;
    q4[1] = (i < nx && j + 1 <= ny) ? Qyk[(size_t)(j + 1) * nx + i] : 0.0;
    q4[2] = (i < nx && j0 <= ny) ? Qyk[(size_t)j0 * nx + i] : 0.0;
    q4[3] = (lane < TY && j0 + lane < ny && i0 + TX <= nx) ? Qxk[(size_t)(j0 + lane) * (nx + 1) + i0 + TX] : 0.0;
}

// Donor-cell face flux F_L = q * c_up of x-face f (row pointers of C and of the exterior values), rules of
// _fct_fluxes (sigma_baroclinic.py:482-491): interior q >= 0 takes the low side; open edges as the WENO faces;
// closed edges 0.
__device__ __forceinline__ double donor_x(const double* __restrict__ crow, const double* __restrict__ erow,
                                          double q, int f, int nx, unsigned flags) {
    if (f > 0 && f < nx) return q * (q >= 0 ? crow[f - 1] : crow[f]);
    if (f == 0) return (flags & OPEN_W) ? q * (q > 0 ? erow[0] : crow[0]) : 0.0;
    return (flags & OPEN_E) ? q * (q < 0 ? erow[nx - 1] : crow[nx - 1]) : 0.0;
}

// Same for y-face f of column i (cp, ep: plane pointers).
__device__ __forceinline__ double donor_y(const double* __restrict__ cp, const double* __restrict__ ep, double q,
                                          int f, int i, int ny, int nx, unsigned flags) {
    if (f > 0 && f < ny) return q * (q >= 0 ? cp[(size_t)(f - 1) * nx + i] : cp[(size_t)f * nx + i]);
    if (f == 0) return (flags & OPEN_S) ? q * (q > 0 ? ep[i] : cp[i]) : 0.0;
    return (flags & OPEN_N) ? q * (q < 0 ? ep[(size_t)(ny - 1) * nx + i] : cp[(size_t)(ny - 1) * nx + i]) : 0.0;
}

__device__ __forceinline__ double dmax(double a, double b) { return b > a ? b : a; }
__device__ __forceinline__ double dmin(double a, double b) { return b < a ? b : a; }

// ------------------------------------------------------------------------------------------------ faces
// Face fluxes fx, fy of cin on every owned face: WENO (the faces_h helper; _advect_faces_h,
// sigma_baroclinic.py:265-321) or UP3 / UP1 (the x faces of flux_x, tracers.py:194-206).
template <int SCH, int TX, int TY>
__global__ __launch_bounds__(TX* TY) void k_faces(const double* __restrict__ cin, const double* __restrict__ ext,
                                                  const double* __restrict__ Qx, const double* __restrict__ Qy,
                                                  const uint8_t* __restrict__ ok, double* __restrict__ fx,
                                                  double* __restrict__ fy, int nt, int n, int ny, int nx,
                                                  unsigned flags) {
    constexpr int H = 3, LX = TX + 2 * H, LY = TY + 2 * H;
    __shared__ double sc[LY * LX];
    __shared__ uint8_t so[LY * LX];
    __shared__ double sfx[TY * (TX + 1)];
    __shared__ double sfy[(TY + 1) * TX];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
    const int i0 = blockIdx.x * TX, j0 = blockIdx.y * TY;
    const size_t plane = (size_t)ny * nx;
    for (int p = blockIdx.z; p < nt * n; p += gridDim.z) {
        const int k = p % n;
        const size_t pc = (size_t)p * plane, kc = (size_t)k * plane;
        __syncthreads();
        load_tile<TX, TY, H>(sc, so, cin + pc, ok + kc, i0, j0, ny, nx);
        __syncthreads();
        tile_faces<SCH, TX, TY>(sfx, sfy, sc, so, Qx + (size_t)k * ny * (nx + 1), Qy + (size_t)k * (ny + 1) * nx,
                                ext ? ext + pc : nullptr, i0, j0, ny, nx, flags);
        __syncthreads();
        for (int idx = tid; idx < TY * (TX + 1); idx += TX * TY) {
            const int ly = idx / (TX + 1), lf = idx - ly * (TX + 1);
            const int j = j0 + ly, f = i0 + lf;
            if (j < ny && f <= nx && (lf < TX || f == nx)) fx[((size_t)p * ny + j) * (nx + 1) + f] = sfx[idx];
        }
        for (int idx = tid; idx < (TY + 1) * TX; idx += TX * TY) {
            const int lf = idx / TX, lx = idx - lf * TX;
            const int f = j0 + lf, i = i0 + lx;
            if (i < nx && f <= ny && (lf < TY || f == ny)) fy[((size_t)p * (ny + 1) + f) * nx + i] = sfy[idx];
        }
    }
}

// Zalesak ratios (eqs. 7-13) of each ok cell: bounds over the cell, i +- 1, j +- 1 and k +- 1 (ok cells only) of
// C and c_L; P+- from the antidiffusive fluxes on the cell's faces; R = min(1, Q/P) or 0
// (_fct_coeff_kernel, sigma_baroclinic.py:413-431).  Cells not ok get R+- = 0.
template <int TX, int TY>
__global__ __launch_bounds__(TX* TY) void k_fct_ratio(const double* __restrict__ C, const double* __restrict__ cL,
                                                      const double* __restrict__ V1, const double* __restrict__ ax,
                                                      const double* __restrict__ ay, const uint8_t* __restrict__ ok,
                                                      double* __restrict__ r_in, double* __restrict__ r_out, int nt,
                                                      int n, int ny, int nx, double dt) {
    const int i = blockIdx.x * TX + threadIdx.x, j = blockIdx.y * TY + threadIdx.y;
    if (i >= nx || j >= ny) return;
    const size_t plane = (size_t)ny * nx;
    for (int p = blockIdx.z; p < nt * n; p += gridDim.z) {
        const int t = p / n, k = p - t * n;
        const size_t xc = (size_t)j * nx + i, x = (size_t)p * plane + xc;
        double ri = 0.0, ro = 0.0;
        if (ok[(size_t)k * plane + xc]) {
            double cmax = -INFINITY, cmin = INFINITY;
            const int dk[7] = {0, 0, 0, 0, 0, -1, 1}, dj[7] = {0, 0, 0, -1, 1, 0, 0}, di[7] = {0, -1, 1, 0, 0, 0, 0};
#pragma unroll
            for (int m = 0; m < 7; ++m) {
                const int kk = k + dk[m], jj = j + dj[m], ii = i + di[m];
                if (kk >= 0 && kk < n && jj >= 0 && jj < ny && ii >= 0 && ii < nx) {
                    const size_t xn = (size_t)jj * nx + ii;
                    if (ok[(size_t)kk * plane + xn]) {
                        const size_t g = ((size_t)t * n + kk) * plane + xn;
                        const double a = C[g], b = cL[g];
                        cmax = dmax(dmax(cmax, a), b);
                        cmin = dmin(dmin(cmin, a), b);
                    }
                }
            }
            const size_t gx = ((size_t)p * ny + j) * (nx + 1) + i;
            const size_t gy = ((size_t)p * (ny + 1) + j) * nx + i;
            const double aw = ax[gx], ae = ax[gx + 1], as = ay[gy], an = ay[gy + nx];
            const double p_in = dt * (dmax(aw, 0.0) - dmin(ae, 0.0) + dmax(as, 0.0) - dmin(an, 0.0));
            const double p_out = dt * (dmax(ae, 0.0) - dmin(aw, 0.0) + dmax(an, 0.0) - dmin(as, 0.0));
            const double v1 = V1[(size_t)k * plane + xc], cl = cL[x];
            const double q_in = dmax(v1 * (cmax - cl), 0.0);
            const double q_out = dmax(v1 * (cl - cmin), 0.0);
            ri = p_in > 0.0 ? dmin(1.0, q_in / p_in) : 0.0;
            ro = p_out > 0.0 ? dmin(1.0, q_out / p_out) : 0.0;
        }
        r_in[x] = ri;
        r_out[x] = ro;
    }
}

// Face coefficients from the ratios of the two adjacent cells (_fct_coeff_kernel :436-447): x-face f (interior)
// C = min(R_in[f], R_out[f-1]) if A >= 0 else min(R_in[f-1], R_out[f]); edge faces 0.
__device__ __forceinline__ double coeff(double a, double rin_hi, double rout_lo, double rin_lo, double rout_hi) {
    return a >= 0.0 ? dmin(rin_hi, rout_lo) : dmin(rin_lo, rout_hi);
}

// Interface kf's z face (1 <= kf <= n-1, between layers kf-1 and kf) from a column window w[0..5] = c[kf-3 .. kf+2]
// (each layer's own ok in b; false and value 0 outside the column): the WENO face times q, upwind-ordered from the
// sign of q as along x and y (_advect_div's vertical pass, sigma_baroclinic.py:335-353), or the UP3 / UP1 / centred
// face value, not yet times q (_up3_faces, _centred_faces along the layer index, rhs_numpy tracers.py:148-150).
// known: every value of the window is 0 or of magnitude in [2^-400, 2^100] (face_value's fast range).
template <int SCH>
__device__ __forceinline__ double zface_w(const double* w, const bool* b, double q, bool known = false) {
    if constexpr (SCH == SCH_C6) {
        const double a = w[2] + w[3];
        const bool m4 = b[1] && b[2] && b[3] && b[4];
        if (m4 && b[0] && b[5]) return (37.0 * a - 8.0 * (w[1] + w[4]) + (w[0] + w[5])) / 60.0;
        if (m4) return (7.0 * a - (w[1] + w[4])) / 12.0;
        return 0.5 * a;
    } else if constexpr (SCH == SCH_UP1) {
        return q >= 0.0 ? w[2] : w[3];
    } else if constexpr (SCH == SCH_UP3) {
        if (q >= 0.0) return (b[1] && b[2] && b[3]) ? (-w[1] + 5.0 * w[2] + 2.0 * w[3]) / 6.0 : w[2];
        return (b[2] && b[3] && b[4]) ? (2.0 * w[2] + 5.0 * w[3] - w[4]) / 6.0 : w[3];
    } else {
        if (q >= 0.0)
            return q * face_value<SCH == SCH_Z>(w[0], w[1], w[2], w[3], w[4], b[0], b[1], b[2], b[3], b[4], known);
        return q * face_value<SCH == SCH_Z>(w[5], w[4], w[3], w[2], w[1], b[5], b[4], b[3], b[2], b[1], known);
    }
}

// waves per SIMD the RK stage kernel is compiled for (its register budget: 512 / waves VGPRs)
#ifndef OCX_RK_WAVES
#define OCX_RK_WAVES 3
#endif

template <int STAGE, int SCH, bool VERT, int TX, int TY>
__global__ __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(OCX_RK_WAVES))) void k_rk_stage(const double* __restrict__ cin, const double* __restrict__ C,
                                                     const double* __restrict__ ext, const double* __restrict__ V0,
                                                     const double* __restrict__ V1, const double* __restrict__ Qx,
                                                     const double* __restrict__ Qy, const double* __restrict__ Fz,
                                                     const uint8_t* __restrict__ ok, double* __restrict__ q,
                                                     double* __restrict__ cout, double* __restrict__ fxs,
                                                     int flux_face_i, int nt, int n, int ny, int nx, double dt,
                                                     unsigned flags) {
    // One block marches one tile of one tracer up its layers (1-D grid of nt * gx * gy blocks, XCD-contiguous):
    // each layer's x and y faces go through the LDS tile; each thread keeps its column's window c[k-2 .. k+3] in
    // registers and forms the z face between layers k and k+1 once, carrying it as the next layer's bottom face.
    constexpr bool ZF = VERT || SCH >= SCH_UP3;
    constexpr int H = 3, LX = TX + 2 * H, LY = TY + 2 * H;
    __shared__ double sc[LY * LX];
    __shared__ uint8_t so[LY * LX];
    __shared__ double sfx[TY * (TX + 1)];
    __shared__ double sfy[(TY + 1) * TX];
    // per layer parity: set when a value of that layer's tile lies outside the WENO divisions' fast range
    constexpr bool WENO = SCH <= SCH_Z;
    __shared__ int sbad[2];
    // the thread's cell in the tile: row ty, column tx (one row per wavefront for 64-wide tiles, where the row is
    // the wave index itself)
    const int tid = threadIdx.y * 64 + threadIdx.x;
    const int ty = TX == 64 ? (int)threadIdx.y : tid / TX, tx = TX == 64 ? (int)threadIdx.x : tid - ty * TX;
    const int gx = (nx + TX - 1) / TX, gy = (ny + TY - 1) / TY;
    const long long nb = (long long)gx * gy * nt;
    // tracer fastest: the nt blocks of a tile run together on one XCD and share its per-layer fields through L2
    long long L = xcd_logical((long long)blockIdx.x, nb);
    const int t = (int)(L % nt);
    L /= nt;
    const int bx = (int)(L % gx);
    const int by = (int)(L / gx);
    const int i0 = bx * TX, j0 = by * TY;
    const size_t plane = (size_t)ny * nx;
    const int j = j0 + ty, i = i0 + tx;
    const bool cell = j < ny && i < nx;
    const size_t xc = cell ? (size_t)j * nx + i : 0;
    // every face of the tile is an interior face of the domain (block-uniform): those tiles take a copy of the face
    // code with no edge rules and no bounds tests (39.4 -> 38.3 ms for the sigma RK3 step at 1024^2 x 50 x 10)
    const bool inner = i0 > 0 && i0 + TX < nx && j0 > 0 && j0 + TY < ny;
    const double* ccol = cin + (size_t)t * n * plane + xc;
    double w[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    bool b[6] = {false, false, false, false, false, false};
    auto wload = [&](int m, int kk) {
        const bool in = cell && kk >= 0 && kk < n;
        w[m] = in ? ccol[(size_t)kk * plane] : 0.0;
        b[m] = in && ok[(size_t)kk * plane + xc] != 0;
    };
    // bit m: w[m] is 0 or of magnitude in [2^-400, 2^100]
    unsigned wfast = 0;
    if constexpr (ZF) {
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            wload(m, m - 2);
            if constexpr (WENO) wfast |= (unsigned)fast_value(w[m]) << m;
        }
    }
    // the bottom face of layer 0: 0 for WENO, Fz[0] * 0 for the rhs_numpy faces
    double fb = 0.0;
    if constexpr (SCH >= SCH_UP3) fb = cell ? Fz[xc] * 0.0 : 0.0;
    // the tile of layer k (with its halo) is fetched into registers one layer ahead and stored to LDS at the top of
    // the layer's iteration, so its loads are in flight while the previous layer's faces are formed
    constexpr int NT = TX * TY, NL = (LY * LX + NT - 1) / NT;
    double pv[NL];
    uint8_t po[NL];
    bool pfast = true;
    auto fetch = [&](int kk) {
        const double* src = cin + ((size_t)t * n + kk) * plane;
        const uint8_t* okp = ok + (size_t)kk * plane;
#pragma unroll
        for (int m = 0; m < NL; ++m) {
            const int idx = tid + m * NT;
            const int ly = idx / LX, lx = idx - ly * LX;
            const int jj = j0 - H + ly, ii = i0 - H + lx;
            const bool in = idx < LY * LX && jj >= 0 && jj < ny && ii >= 0 && ii < nx;
            pv[m] = in ? src[(size_t)jj * nx + ii] : 0.0;
            po[m] = in ? okp[(size_t)jj * nx + ii] : (uint8_t)0;
        }
    };
    auto check = [&]() {
        pfast = true;
#pragma unroll
        for (int m = 0; m < NL; ++m) pfast = pfast && fast_value(pv[m]);
    };
    // the face transports of the thread's face items, likewise one layer ahead
    double qn4[4], qc4[4];
    auto fetch_q = [&](int kk) {
        tile_face_q_waves<TX, TY>(qn4, Qx + (size_t)kk * ny * (nx + 1), Qy + (size_t)kk * (ny + 1) * nx, i0, j0, ny,
                                  nx);
    };
    fetch(0);
    fetch_q(0);
    if constexpr (WENO) {
        if (tid < 2) sbad[tid] = 0;
        __syncthreads();
    }
    for (int k = 0; k < n; ++k) {
        const int p = t * n + k;
        const size_t pc = (size_t)p * plane, kc = (size_t)k * plane;
        // no barrier needed before the store: the tile (sc, so) was last read by the previous layer's face phase,
        // which a barrier closes; the faces (sfx, sfy) are rewritten only after the barrier below
#pragma unroll
        for (int m = 0; m < NL; ++m) {
            const int idx = tid + m * NT;
            if (idx < LY * LX) {
                sc[idx] = pv[m];
                so[idx] = po[m];
            }
        }
        if constexpr (WENO) {
            check();
            if (__builtin_amdgcn_ballot_w64(!pfast) != 0 && threadIdx.x == 0) sbad[k & 1] = 1;
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) qc4[m] = qn4[m];
        if (k + 1 < n) {
            fetch(k + 1);
            fetch_q(k + 1);
        }
        // the cell's update inputs, issued before the face arithmetic so that their loads overlap it
        const size_t x = pc + xc, xv = kc + xc;
        double c0 = 0.0, v0 = 0.0, v1 = 0.0, qo = 0.0, qz = 0.0;
        if (cell) {
            if constexpr (STAGE != 0) {
                c0 = C[x];
                v0 = V0[xv];
                v1 = V1[xv];
            }
            if constexpr (STAGE == 2 || STAGE == 3) qo = q[x];
            if constexpr (ZF) {
                if (k + 1 < n) qz = Fz[(size_t)(k + 1) * plane + xc];
            }
        }
        __syncthreads();
        // the flag of layer k is read below; the other parity's is cleared for layer k + 1 (its readers of layer
        // k - 1 passed the previous barrier, its writers of layer k + 1 come after the next one)
        bool known = false;
        if constexpr (WENO) {
            known = sbad[k & 1] == 0;
            if (tid == 0) sbad[(k + 1) & 1] = 0;
        }
        if (inner)
            tile_faces_waves<SCH, TX, TY, true>(sfx, sfy, sc, so, ext ? ext + pc : nullptr, i0, j0, ny, nx, flags, k,
                                                qc4, known);
        else
            tile_faces_waves<SCH, TX, TY>(sfx, sfy, sc, so, ext ? ext + pc : nullptr, i0, j0, ny, nx, flags, k, qc4,
                                          known);
        __syncthreads();
        double ft = 0.0;
        if constexpr (ZF) {
            // window w = c[k-2 .. k+3] = c[kf-3 .. kf+2] of the interface kf = k+1
            if (cell && k + 1 < n) {
                if constexpr (SCH >= SCH_UP3)
                    ft = qz * zface_w<SCH>(w, b, qz);
                else if (qz != 0.0)
                    ft = zface_w<SCH>(w, b, qz, wfast == 0x3fu);
            }
        }
        if (cell) {
            double d = sfx[ty * (TX + 1) + tx + 1] - sfx[ty * (TX + 1) + tx] + sfy[(ty + 1) * TX + tx] -
                       sfy[ty * TX + tx];
            if constexpr (ZF) d = d + ft - fb;
            const double r = -d;
            if (fxs) {
                if (flux_face_i == i) fxs[(size_t)p * ny + j] = sfx[ty * (TX + 1) + tx];
                if (flux_face_i == nx && i == nx - 1) fxs[(size_t)p * ny + j] = sfx[ty * (TX + 1) + tx + 1];
            }
            if (STAGE == 0) {
                cout[x] = r;
            } else {
                const double q0 = v0 * c0, dV = v1 - v0;
                double qn, V;
                if (STAGE == 1) {
                    qn = q0 + dt * r;
                    V = v0 + dV;
                    q[x] = qn;
                } else if (STAGE == 2) {
                    qn = 0.75 * q0 + 0.25 * (qo + dt * r);
                    V = v0 + 0.5 * dV;
                    q[x] = qn;
                } else if (STAGE == 3) {
                    qn = q0 / 3.0 + 2.0 / 3.0 * (qo + dt * r);
                    V = v1;
                } else if (STAGE == 4) {
                    qn = q0 + 0.5 * dt * r;
                    V = v0 + 0.5 * dV;
                } else {
                    qn = q0 + dt * r;
                    V = v1;
                }
                cout[x] = V > 0 ? qn / V : c0;
            }
        }
        if constexpr (ZF) {
            fb = ft;
#pragma unroll
            for (int m = 0; m < 5; ++m) {
                w[m] = w[m + 1];
                b[m] = b[m + 1];
            }
            wload(5, k + 4);
            if constexpr (WENO) wfast = (wfast >> 1) | ((unsigned)fast_value(w[5]) << 5);
        }
    }
}

// (F(c0) + F(c1))/6 + 2 F(c2)/3 of the three stages' face columns (tracers.py:375).
__global__ void k_rk_face_column(const double* __restrict__ f0, const double* __restrict__ f1,
                                 const double* __restrict__ f2, double* __restrict__ out, size_t count) {
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < count) out[x] = (f0[x] + f1[x]) / 6.0 + 2.0 * f2[x] / 3.0;
}

// Face values along one axis of an array viewed as (outer, n, inner) cells with (outer, nf, inner) faces: face f
// lies between cells f-1 and f of its line, faces 1 .. n-1 take the scheme's value (WENO5-Z/JS upwind-ordered by
// the sign of flux, flux >= 0 taking the low side; UP3 / UP1 likewise; centred whatever the flux), every other
// face 0.  Cells outside 0 .. n-1 read value 0 and are not valid.  One thread per face; consecutive threads take
// consecutive inner indices (or, with inner = 1, consecutive faces), so the loads are coalesced either way.
// _weno5_faces / _weno5_rows, _up3_faces, _centred_faces (core/dynamics/sigma_baroclinic.py:236-262, 564-578,
// 164-187, 632-652) as TracerAdvection.faces calls them (core/components/tracers.py:110-115).
template <int SCH>
__global__ __launch_bounds__(256) void k_line_faces(const double* __restrict__ c, const double* __restrict__ flux,
                                                    const uint8_t* __restrict__ ok, long long outer, int n, int nf,
                                                    long long inner, double* __restrict__ out) {
    const long long total = outer * nf * inner;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long r = e % inner;
        const long long q = e / inner;
        const int f = (int)(q % nf);
        const long long o = q / nf;
        if (f < 1 || f > n - 1) {
            out[e] = 0.0;
            continue;
        }
        const long long base = o * n * inner + r;
        double u[6];
        uint8_t v[6];
#pragma unroll
        for (int m = 0; m < 6; ++m) {
            const int cell = f - 3 + m;
            const bool in = cell >= 0 && cell < n;
            const long long x = base + (long long)(in ? cell : 0) * inner;
            u[m] = in ? c[x] : 0.0;
            v[m] = in && ok[x] != 0;
        }
        const double qf = SCH == SCH_C6 ? 0.0 : flux[e];
        double val;
        if (SCH == SCH_C6) val = centred_face(u, v, 1, 3);
        else if (SCH == SCH_UP3) val = up_face<true>(u, v, 1, 3, qf);
        else if (SCH == SCH_UP1) val = up_face<false>(u, v, 1, 3, qf);
        else val = lds_face<SCH == SCH_Z>(u, v, 1, 3, qf);
        out[e] = val;
    }
}

// Open-edge face values of one horizontal axis (TracerAdvection.open_faces, core/components/tracers.py:117-137), in
// place on face (nt, nz, ny, nx+1) (axis 0) or (nt, nz, ny+1, nx) (axis 1): on each open edge of the axis the
// exterior value ext where the flux Q (nz, ny, nx+1) / (nz, ny+1, nx) enters the domain (Q > 0 on the west / south
// edge, Q < 0 on the east / north edge), the boundary cell of c otherwise (also where Q is 0 or NaN).
__global__ void k_open_faces(double* __restrict__ face, const double* __restrict__ c, const double* __restrict__ Q,
                             const double* __restrict__ ext, int nt, int nz, int ny, int nx, int axis, int lo_open,
                             int hi_open) {
    const int m = axis == 0 ? ny : nx;                   // faces along one edge of one (t, k) plane
    const long long total = (long long)nt * nz * m * 2;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int side = (int)(e % 2);
        const long long q = e / 2;
        const int a = (int)(q % m);
        const long long p = q / m;                       // plane t * nz + k
        const int k = (int)(p % nz);
        if (side == 0 ? !lo_open : !hi_open) continue;
        long long fidx, cidx, qidx;
        if (axis == 0) {
            const int fi = side == 0 ? 0 : nx, ci = side == 0 ? 0 : nx - 1;
            fidx = (p * ny + a) * (nx + 1) + fi;
            cidx = (p * ny + a) * nx + ci;
            qidx = ((long long)k * ny + a) * (nx + 1) + fi;
        } else {
            const int fj = side == 0 ? 0 : ny, cj = side == 0 ? 0 : ny - 1;
            fidx = (p * (ny + 1) + fj) * nx + a;
            cidx = (p * ny + cj) * nx + a;
            qidx = ((long long)k * (ny + 1) + fj) * nx + a;
        }
        const double sgn = side == 0 ? 1.0 : -1.0;
        face[fidx] = (sgn * Q[qidx] > 0.0) ? ext[cidx] : c[cidx];
    }
}

// Donor-cell face fluxes of every face (TracerAdvection.upwind_faces_h, tracers.py:178-193).
__global__ void k_upwind_faces(const double* __restrict__ c, const double* __restrict__ ext,
                               const double* __restrict__ Qx, const double* __restrict__ Qy, double* __restrict__ fx,
                               double* __restrict__ fy, int nt, int n, int ny, int nx, unsigned flags) {
    const size_t nfx = (size_t)nt * n * ny * (nx + 1), nfy = (size_t)nt * n * (ny + 1) * nx;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < nfx + nfy; g += (size_t)gridDim.x * blockDim.x) {
        if (g < nfx) {
            const int f = (int)(g % (nx + 1));
            const size_t r = g / (nx + 1);
            const int j = (int)(r % ny);
            const size_t p = r / ny;
            const int k = (int)(p % n);
            const double* crow = c + (p * ny + j) * nx;
            const double* erow = (ext ? ext : c) + (p * ny + j) * nx;
            fx[g] = donor_x(crow, erow, Qx[((size_t)k * ny + j) * (nx + 1) + f], f, nx, flags);
        } else {
            const size_t h = g - nfx;
            const int i = (int)(h % nx);
            const size_t r = h / nx;
            const int f = (int)(r % (ny + 1));
            const size_t p = r / (ny + 1);
            const int k = (int)(p % n);
            const double* cp = c + p * ny * nx;
            const double* ep = (ext ? ext : c) + p * ny * nx;
            fy[h] = donor_y(cp, ep, Qy[((size_t)k * (ny + 1) + f) * nx + i], f, i, ny, nx, flags);
        }
    }
}

// Face coefficients Cx, Cy from the ratios (for the coefficient entry point; the step fuses them into
// k_fct_final).
__global__ void k_fct_face_coeff(const double* __restrict__ ax, const double* __restrict__ ay,
                                 const double* __restrict__ r_in, const double* __restrict__ r_out,
                                 double* __restrict__ cx, double* __restrict__ cy, int nt, int n, int ny, int nx) {
    const size_t nfx = (size_t)nt * n * ny * (nx + 1), nfy = (size_t)nt * n * (ny + 1) * nx;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < nfx + nfy; g += (size_t)gridDim.x * blockDim.x) {
        if (g < nfx) {
            const int f = (int)(g % (nx + 1));
            const size_t r = g / (nx + 1);
            const size_t row = r * nx;
            cx[g] = (f > 0 && f < nx) ? coeff(ax[g], r_in[row + f], r_out[row + f - 1], r_in[row + f - 1], r_out[row + f])
                                      : 0.0;
        } else {
            const size_t h = g - nfx;
            const int i = (int)(h % nx);
            const size_t r = h / nx;
            const int f = (int)(r % (ny + 1));
            const size_t p = r / (ny + 1);
            const size_t pc = p * ny * nx;
            cy[h] = (f > 0 && f < ny) ? coeff(ay[h], r_in[pc + (size_t)f * nx + i], r_out[pc + (size_t)(f - 1) * nx + i],
                                              r_in[pc + (size_t)(f - 1) * nx + i], r_out[pc + (size_t)f * nx + i])
                                      : 0.0;
        }
    }
}

// Per-tracer sums of the open-edge face fluxes: sum_W G - sum_E G + sum_S G - sum_N G (tracers.py:65-78).  One
// block per tracer, fixed summation order (deterministic).
__global__ void k_open_flux(const double* __restrict__ gedge, double* __restrict__ out, int nt, int n, int ny, int nx,
                            unsigned flags) {
    __shared__ double red[256];
    const int t = blockIdx.x;
    const size_t nyl = (size_t)nt * n * ny, nxl = (size_t)nt * n * nx;
    const size_t my = (size_t)n * ny, mx = (size_t)n * nx;
    double s = 0.0;
    for (size_t e = threadIdx.x; e < my; e += blockDim.x) {
        const size_t g = (size_t)t * my + e;
        if (flags & OPEN_W) s += gedge[g];
        if (flags & OPEN_E) s -= gedge[nyl + g];
    }
    for (size_t e = threadIdx.x; e < mx; e += blockDim.x) {
        const size_t g = (size_t)t * mx + e;
        if (flags & OPEN_S) s += gedge[2 * nyl + g];
        if (flags & OPEN_N) s -= gedge[2 * nyl + nxl + g];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[t] = red[0];
}

// dst[0..count) = value (a kernel, not hipMemsetAsync, so that it may sit in a captured graph).
__global__ void k_fill(double* __restrict__ dst, double value, size_t count) {
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < count; x += (size_t)gridDim.x * blockDim.x)
        dst[x] = value;
}

// Self-test of the fast divisions: for n operand pairs (x, y), out[0..n) = div_const(x, 6), [n..2n) = div_const(x, 3),
// [2n..3n) = div_nr(x, y), [3n..4n) = qdiv(x, y), [4n..5n) = the IEEE quotient x / y, [5n..6n) = qdiv_r(x,
// recip_of(y)).
__global__ void k_selftest_div(const double* __restrict__ x, const double* __restrict__ y, size_t n,
                               double* __restrict__ out) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const double a = x[e], b = y[e];
    out[e] = div_const(a, 6.0, 1.0 / 6.0);
    out[n + e] = div_const(a, 3.0, 1.0 / 3.0);
    out[2 * n + e] = div_nr(a, b);
    out[3 * n + e] = qdiv(a, b);
    out[4 * n + e] = a / b;
    out[5 * n + e] = qdiv_r(a, recip_of(b));
}

// Numerical mixing of the advection of c (Burchard & Rennau 2008, eq. 24), with c^2 advected in the same stack:
// m = V1 (<c^2> - c^2) / dt per cell, in the reference's operation order (VarianceBudget.advective,
// core/components/variance_budget.py:60-62; mix_num, core/model/driver.py:905); accumulate adds it to out.
__global__ void k_mix_num(const double* __restrict__ c, const double* __restrict__ c2, const double* __restrict__ V1,
                          double dt, int accumulate, double* __restrict__ out, size_t cells) {
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < cells; x += (size_t)gridDim.x * blockDim.x) {
        const double cx = c[x];
        const double m = V1[x] * (c2[x] - cx * cx) / dt;
        out[x] = accumulate ? out[x] + m : m;
    }
}

// Per-tracer content sum_k,j,i V[k,j,i] * c[t,k,j,i] in FP64: compensated (Neumaier) partial sums in a fixed
// order, then fixed-shape trees, so the result is deterministic and its rounding error stays near one ulp of the
// sum.
__device__ __forceinline__ void neumaier_add(double& s, double& comp, double x) {
    const double t = s + x;
    comp += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
}

// Content sums in two passes with a fixed partition, so the result does not depend on the launch: pass 1 gives
// each of CONTENT_PARTS blocks per tracer one contiguous slice of the cells (compensated sum over its threads'
// strided elements, then a fixed tree over the block); pass 2 folds a tracer's CONTENT_PARTS partials the same way.
constexpr int CONTENT_PARTS = 256;

__device__ __forceinline__ void block_fold(double& s, double& comp, double* red, double* rc) {
    red[threadIdx.x] = s;
    rc[threadIdx.x] = comp;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double a = red[threadIdx.x], ca = rc[threadIdx.x] + rc[threadIdx.x + w];
            neumaier_add(a, ca, red[threadIdx.x + w]);
            red[threadIdx.x] = a;
            rc[threadIdx.x] = ca;
        }
        __syncthreads();
    }
    s = red[0];
    comp = rc[0];
}

// grid (CONTENT_PARTS, nt), 256 threads: part[t][b] = (sum, compensation) of slice b of tracer t
__global__ __launch_bounds__(256) void k_content_part(const double* __restrict__ c, const double* __restrict__ V,
                                                      double* __restrict__ part, size_t cells) {
    __shared__ double red[256], rc[256];
    const int t = blockIdx.y, b = blockIdx.x;
    const size_t per = (cells + CONTENT_PARTS - 1) / CONTENT_PARTS;
    const size_t x0 = (size_t)b * per, x1 = x0 + per < cells ? x0 + per : cells;
    const double* ct = c + (size_t)t * cells;
    double s = 0.0, comp = 0.0;
    for (size_t x = x0 + threadIdx.x; x < x1; x += blockDim.x) neumaier_add(s, comp, V[x] * ct[x]);
    block_fold(s, comp, red, rc);
    if (threadIdx.x == 0) {
        part[((size_t)t * CONTENT_PARTS + b) * 2] = s;
        part[((size_t)t * CONTENT_PARTS + b) * 2 + 1] = comp;
    }
}

// grid nt, CONTENT_PARTS threads: out[t] = the folded partials of tracer t
__global__ __launch_bounds__(CONTENT_PARTS) void k_content_fold(const double* __restrict__ part,
                                                                double* __restrict__ out) {
    __shared__ double red[CONTENT_PARTS], rc[CONTENT_PARTS];
    const int t = blockIdx.x;
    double s = part[((size_t)t * CONTENT_PARTS + threadIdx.x) * 2];
    double comp = part[((size_t)t * CONTENT_PARTS + threadIdx.x) * 2 + 1];
    block_fold(s, comp, red, rc);
    if (threadIdx.x == 0) out[t] = s + comp;
}

// Periodic halo (core/model/periodic.py:71-95): gather of ``lead`` slabs of a (sy, sx) array into (dy, dx) with
// source index (pos - halo) mod n on each periodic axis (n = unique cells or faces), or the crop (inverse).
template <typename T>
__global__ void k_pad(const T* __restrict__ src, T* __restrict__ dst, size_t lead, int sy, int sx, int dy, int dx,
                      int halo_y, int halo_x, int ny_unique, int nx_unique, int crop) {
    const size_t total = lead * (size_t)dy * dx;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
        const int ii = (int)(g % dx);
        const size_t r = g / dx;
        const int jj = (int)(r % dy);
        const size_t l = r / dy;
        int si, sj;
        if (crop) {
            si = ii + halo_x;
            sj = jj + halo_y;
        } else {
            si = halo_x ? ((ii - halo_x) % nx_unique + nx_unique) % nx_unique : ii;
            sj = halo_y ? ((jj - halo_y) % ny_unique + ny_unique) % ny_unique : jj;
        }
        dst[g] = src[(l * sy + sj) * (size_t)sx + si];
    }
}

}  // namespace ocx
