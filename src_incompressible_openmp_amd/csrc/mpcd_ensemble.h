/*
 * mpcd_ensemble.h -- batched ensembles: many independent channels (replicas
 * of uneven grid sizes, forcings, particle counts and seeds) advanced
 * together, each stage of the step one launch over all replicas.  Included
 * once, at the end of mpcd_engine.hip, whose kernels and device functions it
 * uses; the C ABI it implements is declared in include/mpcd_gpu.h.
 *
 * Layout.  The replicas' particles are concatenated (replica r owns slots
 * [poff[r], poff[r+1])) and so are their cells (global cell = coff[r] + the
 * replica's own cell), so one counting sort bins every replica at once and
 * keeps each replica's particles in its own range of slots through the
 * kick's reorder.  Particle ids (the key of every draw) and the cell numbers
 * in the draws are the replica's own, and its seeds are the benchmark
 * driver's plus seedKey * SEED_KEY_MIX: replica r advances bit for bit as a
 * single simulation given the same seed key (mpcd_gpu_sim_set_seed_key).
 *
 * Per-particle and per-cell kernels find their replica by a binary search
 * of the offset tables (once per block when the block lies in one replica).
 * beta's sums use blocks aligned to each replica's cells: the same partials
 * and the same tree as a single simulation's.  The zone passes run every
 * replica's zones in one launch of the fused zone kernel (its kEns form
 * reads each zone's inputs from its replica's table).  The per-step values
 * (seeds, grid shifts) are computed on the device from a step counter, so
 * the step is a fixed launch sequence: it is captured into a HIP graph per
 * buffer parity and replayed.
 */

namespace {

constexpr uint64_t SEED_KEY_MIX = 0x9E3779B97F4A7C15ULL;

/* one replica's constants on the device */
struct EnsRep {
    Geom g;
    BCParams b;
    CollideParams cp;
    double xmax, ymax, Vc, K, rhoT, gamma, dtr, betaEOS, blend, betaRepair;
    int useOptimal, anyMask;
    int poff, coff, boff, nbc; /* first slot, first cell, first closure block, closure blocks */
    uint64_t seedOff;
};

/* largest r with off[r] <= i (off: R+1 ascending entries) */
__device__ __forceinline__ int ens_find(const int* __restrict__ off, int R, int i) {
    int lo = 0, hi = R - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

/* The replica of element i, for a block over the elements [b0, b0 + nb):
 * thread 0 searches for b0, and the whole block takes that replica when it
 * holds the block; otherwise each thread searches.  Every thread of the
 * block calls it once, before any return. */
__device__ __forceinline__ int ens_rep(const int* __restrict__ off, int R, int i, int b0, int nb, int* sh) {
    if (threadIdx.x == 0) {
        const int r = ens_find(off, R, b0);
        *sh = (off[r + 1] >= b0 + nb) ? r : -1;
    }
    __syncthreads();
    const int r = *sh;
    return r >= 0 ? r : ens_find(off, R, i);
}

/* per replica: the grid shift, the zone seed bases of both layouts, and the
 * wall and rotation seeds of the step in *step (the benchmark driver's
 * seeds, :240-279, offset by the replica's key) */
__global__ void k_ens_seeds(const EnsRep* __restrict__ reps, int R, const long long* __restrict__ step,
                            double* __restrict__ shift, uint64_t* __restrict__ zseed, uint64_t* __restrict__ seeds) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const uint64_t us = (uint64_t)*step, off = reps[r].seedOff;
    double sx, sy;
    mpcd_draw_shift(130363ULL + 10007ULL * us + off, reps[r].g.a0, &sx, &sy);
    shift[2 * r] = sx;
    shift[2 * r + 1] = sy;
    /* the next step's shift and wall seed, for a kick that streams it */
    mpcd_draw_shift(130363ULL + 10007ULL * (us + 1ULL) + off, reps[r].g.a0, &sx, &sy);
    shift[2 * (size_t)R + 2 * r] = sx;
    shift[2 * (size_t)R + 2 * r + 1] = sy;
    seeds[2 * (size_t)R + r] = 104729ULL + 10007ULL * (us + 1ULL) + off;
    const uint64_t zs = 88172645463325252ULL + 10007ULL * us + off;
    zseed[r] = zs;
    zseed[R + r] = zs + 4099ULL;
    seeds[2 * r] = 104729ULL + 10007ULL * us + off;
    seeds[2 * r + 1] = 433494437ULL + 10007ULL * us + off;
}

__global__ void k_ens_step_end(long long* step) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

__global__ void k_ens_set_step(long long* step, long long s) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step = s;
}

/* force + stream + wall conditions (k_stream_bc per replica) and both global
 * cell ids of every particle */
__global__ void __launch_bounds__(TPB) k_ens_stream(const EnsRep* __restrict__ reps, const int* __restrict__ poff,
                                                    int R, int N, const uint64_t* __restrict__ seeds,
                                                    const double* __restrict__ shift, double* __restrict__ x,
                                                    double* __restrict__ y, double* __restrict__ vx,
                                                    double* __restrict__ vy, const int* __restrict__ pid,
                                                    int* __restrict__ cidS, int* __restrict__ cidU,
                                                    int* __restrict__ err) {
    __shared__ int sh;
    const int b0 = blockIdx.x * TPB, i = b0 + (int)threadIdx.x;
    const int r = ens_rep(poff, R, i < N ? i : N - 1, b0, min(TPB, N - b0), &sh);
    if (i >= N) return;
    const EnsRep& e = reps[r];
    double px = x[i], py = y[i], pvx = vx[i], pvy = vy[i];
    WallAcc acc;
    wall_acc_zero(acc);
    const int rc = bc_particle(e.g, e.b, seeds[2 * r], (uint32_t)pid[i], 1, px, py, pvx, pvy, acc);
    if (rc != 0) atomicOr(err, 1);
    x[i] = px;
    y[i] = py;
    vx[i] = pvx;
    vy[i] = pvy;
    cidS[i] = e.coff + shifted_cell_of(e.g, px, py, shift[2 * r], shift[2 * r + 1], e.xmax, e.ymax);
    cidU[i] = e.coff + cell_of(e.g, px, py);
}

/* SRD collision of every cell (k_collide per replica: the cell number in the
 * draws is the replica's own) */
__global__ void __launch_bounds__(CELL_BLK) k_ens_collide(const EnsRep* __restrict__ reps,
                                                          const int* __restrict__ coff, int R,
                                                          const uint64_t* __restrict__ seeds,
                                                          const int* __restrict__ start, const int* __restrict__ list,
                                                          int C, double* __restrict__ vx, double* __restrict__ vy,
                                                          int cap) {
    extern __shared__ double sv[];
    __shared__ int sh;
    const int c0 = blockIdx.x * CELL_BLK;
    const int c1 = min(c0 + CELL_BLK, C);
    const int c = c0 + (int)threadIdx.x;
    const int r = ens_rep(coff, R, c < C ? c : C - 1, c0, c1 - c0, &sh);
    const int b0 = start[c0], b1 = start[c1];
    const EnsRep& e = reps[r];
    if (b1 - b0 <= cap) {
        double* svx = sv;
        double* svy = sv + cap;
        stage_in(list, b0, b1 - b0, vx, vy, svx, svy);
        __syncthreads();
        if (c < c1) {
            const int o = start[c] - b0;
            collide_cell(e.cp, seeds[2 * r + 1], c - e.coff, nullptr, start[c + 1] - start[c], svx + o, svy + o);
        }
        __syncthreads();
        stage_out(list, b0, b1 - b0, vx, vy, svx, svy);
        return;
    }
    if (c >= c1) return;
    collide_cell(e.cp, seeds[2 * r + 1], c - e.coff, list + start[c], start[c + 1] - start[c], vx, vy);
}

/* cell fields of every cell (k_cell_fields per replica) */
__global__ void __launch_bounds__(CELL_BLK) k_ens_cell_fields(const EnsRep* __restrict__ reps,
                                                              const int* __restrict__ coff, int R,
                                                              const int* __restrict__ start,
                                                              const int* __restrict__ list, int C,
                                                              const double* __restrict__ vx,
                                                              const double* __restrict__ vy, FieldsOut f, int cap) {
    extern __shared__ double sv[];
    __shared__ int sh;
    const int c0 = blockIdx.x * CELL_BLK;
    const int c1 = min(c0 + CELL_BLK, C);
    const int c = c0 + (int)threadIdx.x;
    const int r = ens_rep(coff, R, c < C ? c : C - 1, c0, c1 - c0, &sh);
    const int b0 = start[c0], b1 = start[c1];
    const EnsRep& e = reps[r];
    CellStats s;
    if (b1 - b0 <= cap) {
        double* svx = sv;
        double* svy = sv + cap;
        stage_in(list, b0, b1 - b0, vx, vy, svx, svy);
        __syncthreads();
        if (c >= c1) return;
        const int o = start[c] - b0;
        s = cell_stats(nullptr, start[c + 1] - start[c], svx + o, svy + o, e.Vc, e.K, e.rhoT);
    } else {
        if (c >= c1) return;
        s = cell_stats(list + start[c], start[c + 1] - start[c], vx, vy, e.Vc, e.K, e.rhoT);
    }
    if (f.N) f.N[c] = s.N;
    if (f.Ux) f.Ux[c] = s.Ux;
    if (f.Uy) f.Uy[c] = s.Uy;
    if (f.Px) f.Px[c] = s.Px;
    if (f.Py) f.Py[c] = s.Py;
    if (f.T) f.T[c] = s.T;
    if (f.rho) f.rho[c] = s.rho;
    if (f.Pkin) f.Pkin[c] = s.Pkin;
    if (f.Pvir) f.Pvir[c] = s.Pvir;
    if (f.P) f.P[c] = s.P;
}

/* particles of each zone of every replica (k_zone_sizes on the replica's
 * cells); a zone too large for the 16-bit working set is flagged (bit 8) */
__global__ void __launch_bounds__(64) k_ens_zone_sizes(const ZoneRect* __restrict__ rects,
                                                       const int* __restrict__ zrep,
                                                       const EnsRep* __restrict__ reps, int nZ,
                                                       const int* __restrict__ start, int* __restrict__ sizes,
                                                       int* __restrict__ err) {
    const int zi = blockIdx.x, lane = threadIdx.x;
    if (zi > nZ) return;
    if (zi == nZ) {
        if (lane == 0) sizes[zi] = 0;
        return;
    }
    const ZoneRect rc = rects[zi];
    const EnsRep& e = reps[zrep[zi]];
    const int* st = start + e.coff;
    const int Ny = e.g.Ny;
    int P = 0;
    for (int gx = rc.xs + lane; gx < rc.xe; gx += 64) P += st[(rc.ye - 1 + Ny * gx) + 1] - st[rc.ys + Ny * gx];
    for (int o = 32; o > 0; o >>= 1) P += __shfl_xor(P, o, 64);
    if (lane == 0) {
        sizes[zi] = P;
        if (P > 65534) atomicOr(err, 8);
    }
}

/* closure repair per replica in blocks aligned to its cells (block b of the
 * launch is block b - boff[r] of replica r): the same per-block partials of
 * (num, den) as k_closure_repair on the replica alone */
__global__ void __launch_bounds__(TPB) k_ens_closure_repair(const EnsRep* __restrict__ reps,
                                                            const int* __restrict__ boff, int R, ClosureArgs a) {
    __shared__ int sh;
    __shared__ double s0[TPB], s1[TPB];
    const int b = blockIdx.x;
    if (threadIdx.x == 0) sh = ens_find(boff, R, b);
    __syncthreads();
    const EnsRep& e = reps[sh];
    const int cl = (b - e.boff) * TPB + (int)threadIdx.x;
    double num = 0.0, den = 0.0;
    if (cl < e.g.Nc) {
        const int c = e.coff + cl;
        const double ax = (a.refPx[c] - a.redPx[c]) / e.gamma;
        const double ay = (a.refPy[c] - a.redPy[c]) / e.gamma;
        a.duRX[c] = ax;
        a.duRY[c] = ay;
        const double ex = a.redUx[c] - a.refUx[c], ey = a.redUy[c] - a.refUy[c];
        num = -(ex * ax + ey * ay);
        den = ax * ax + ay * ay;
    }
    s0[threadIdx.x] = num;
    s1[threadIdx.x] = den;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < (unsigned)s) {
            s0[threadIdx.x] += s0[threadIdx.x + s];
            s1[threadIdx.x] += s1[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.part[2 * b] = s0[0];
        a.part[2 * b + 1] = s1[0];
    }
}

/* beta of every replica: block r reduces replica r's partials (k_reduce_beta) */
__global__ void __launch_bounds__(TPB) k_ens_reduce_beta(const EnsRep* __restrict__ reps,
                                                         const double* __restrict__ part, double* __restrict__ scal) {
    const EnsRep& e = reps[blockIdx.x];
    double* sc = scal + 4 * (size_t)blockIdx.x;
    reduce_parts_block(part + 2 * (size_t)e.boff, e.nbc, 2, sc);
    if (threadIdx.x == 0) beta_from_sums(sc, e.useOptimal, e.betaRepair);
}

/* pdrive with the interzone cross smoothing (k_pdrive per replica) */
__global__ void __launch_bounds__(TPB) k_ens_pdrive(const EnsRep* __restrict__ reps, const int* __restrict__ coff,
                                                    int R, int C, const double* __restrict__ raw,
                                                    const int* __restrict__ mask, double* __restrict__ pd) {
    __shared__ int sh;
    const int b0 = blockIdx.x * TPB, c = b0 + (int)threadIdx.x;
    const int r = ens_rep(coff, R, c < C ? c : C - 1, b0, min(TPB, C - b0), &sh);
    if (c >= C) return;
    const EnsRep& e = reps[r];
    const Geom& g = e.g;
    const double* rw = raw + e.coff;
    const int cl = c - e.coff;
    double v = rw[cl];
    if (e.anyMask && mask[c] != 0) {
        const int ix = cl / g.Ny, iy = cl - ix * g.Ny;
        double acc = rw[cl];
        int n = 1;
        const int ddx[4] = {-1, 1, 0, 0}, ddy[4] = {0, 0, -1, 1};
        for (int k = 0; k < 4; ++k) {
            int jx = ix + ddx[k], jy = iy + ddy[k];
            if (g.perx) jx = wrapi(jx, g.Nx);
            else if (jx < 0 || jx >= g.Nx) continue;
            if (g.pery) jy = wrapi(jy, g.Ny);
            else if (jy < 0 || jy >= g.Ny) continue;
            acc += rw[jy + g.Ny * jx];
            ++n;
        }
        const double ps = acc / (double)n;
        v = (1.0 - e.blend) * rw[cl] + e.blend * ps;
    }
    pd[c] = v;
}

/* duEOS (k_eos per replica) */
__global__ void __launch_bounds__(TPB) k_ens_eos(const EnsRep* __restrict__ reps, const int* __restrict__ coff, int R,
                                                 int C, const double* __restrict__ pdr, double* __restrict__ exo,
                                                 double* __restrict__ eyo) {
    __shared__ int sh;
    const int b0 = blockIdx.x * TPB, c = b0 + (int)threadIdx.x;
    const int r = ens_rep(coff, R, c < C ? c : C - 1, b0, min(TPB, C - b0), &sh);
    if (c >= C) return;
    const EnsRep& e = reps[r];
    const Geom& g = e.g;
    const double* p = pdr + e.coff;
    const int cl = c - e.coff;
    const int ix = cl / g.Ny, iy = cl - ix * g.Ny;
    const double dx = g.dxc, dy = g.dyc;
    double gx, gy;
    if (g.perx) gx = (p[iy + g.Ny * wrapi(ix + 1, g.Nx)] - p[iy + g.Ny * wrapi(ix - 1, g.Nx)]) / (2.0 * dx);
    else if (ix == 0) gx = (p[iy + g.Ny * (ix + 1)] - p[cl]) / dx;
    else if (ix == g.Nx - 1) gx = (p[cl] - p[iy + g.Ny * (ix - 1)]) / dx;
    else gx = (p[iy + g.Ny * (ix + 1)] - p[iy + g.Ny * (ix - 1)]) / (2.0 * dx);
    if (g.pery) gy = (p[wrapi(iy + 1, g.Ny) + g.Ny * ix] - p[wrapi(iy - 1, g.Ny) + g.Ny * ix]) / (2.0 * dy);
    else if (iy == 0) gy = (p[(iy + 1) + g.Ny * ix] - p[cl]) / dy;
    else if (iy == g.Ny - 1) gy = (p[cl] - p[(iy - 1) + g.Ny * ix]) / dy;
    else gy = (p[(iy + 1) + g.Ny * ix] - p[(iy - 1) + g.Ny * ix]) / (2.0 * dy);
    exo[c] = -e.dtr * gx * e.betaEOS;
    eyo[c] = -e.dtr * gy * e.betaEOS;
}

/* the closure's kick and the reorder into red-cell order (k_kick_reorder
 * per replica; each replica's slots stay in its own range) */
__global__ void __launch_bounds__(TPB) k_ens_kick_reorder(
    const EnsRep* __restrict__ reps, const int* __restrict__ poff, int R, int N, const int* __restrict__ list,
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ vx,
    const double* __restrict__ vy, const int* __restrict__ pid, double* __restrict__ xo, double* __restrict__ yo,
    double* __restrict__ vxo, double* __restrict__ vyo, int* __restrict__ pido, const double* __restrict__ duRX,
    const double* __restrict__ duRY, const double* __restrict__ duEX, const double* __restrict__ duEY,
    const double* __restrict__ scal) {
    __shared__ int sh;
    const int b0 = blockIdx.x * TPB, s = b0 + (int)threadIdx.x;
    const int r = ens_rep(poff, R, s < N ? s : N - 1, b0, min(TPB, N - b0), &sh);
    if (s >= N) return;
    const EnsRep& e = reps[r];
    const double beta = scal[4 * (size_t)r + 3];
    const int i = list[s];
    const double px = x[i], py = y[i];
    const int c = e.coff + cell_of(e.g, px, py);
    const double kx = beta * duRX[c], ky = beta * duRY[c];
    double nx, ny;
    kick_velocity(vx[i], vy[i], kx, ky, duEX[c], duEY[c], nx, ny);
    xo[s] = px;
    yo[s] = py;
    vxo[s] = nx;
    vyo[s] = ny;
    pido[s] = pid[i];
}

/* the kick and reorder fused with the next step's force, stream and wall
 * conditions (k_kick_reorder_stream per replica): the state is written
 * once, already streamed, with both global cell ids of the next step and
 * their histograms (cntS, cntU, cleared before) */
__global__ void __launch_bounds__(TPB) k_ens_kick_reorder_stream(
    const EnsRep* __restrict__ reps, const int* __restrict__ poff, int R, int N, const uint64_t* __restrict__ bcSeedN,
    const double* __restrict__ shiftN, const int* __restrict__ list, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ vx, const double* __restrict__ vy,
    const int* __restrict__ pid, double* __restrict__ xo, double* __restrict__ yo, double* __restrict__ vxo,
    double* __restrict__ vyo, int* __restrict__ pido, const double* __restrict__ duRX,
    const double* __restrict__ duRY, const double* __restrict__ duEX, const double* __restrict__ duEY,
    const double* __restrict__ scal, int* __restrict__ cidS, int* __restrict__ cidU, int* __restrict__ cntS,
    int* __restrict__ cntU, int* __restrict__ err) {
    __shared__ int sh;
    const int b0 = blockIdx.x * TPB, s = b0 + (int)threadIdx.x;
    const int r = ens_rep(poff, R, s < N ? s : N - 1, b0, min(TPB, N - b0), &sh);
    if (s >= N) return;
    const int lane = __lane_id();
    const EnsRep& e = reps[r];
    const double beta = scal[4 * (size_t)r + 3];
    const int i = list[s];
    double px = x[i], py = y[i];
    const int c = e.coff + cell_of(e.g, px, py);
    const double kx = beta * duRX[c], ky = beta * duRY[c];
    double nx, ny;
    kick_velocity(vx[i], vy[i], kx, ky, duEX[c], duEY[c], nx, ny);
    const int pi = pid[i];
    WallAcc acc;
    wall_acc_zero(acc);
    const int rc = bc_particle(e.g, e.b, bcSeedN[r], (uint32_t)pi, 1, px, py, nx, ny, acc);
    if (rc != 0) atomicOr(err, 1);
    xo[s] = px;
    yo[s] = py;
    vxo[s] = nx;
    vyo[s] = ny;
    pido[s] = pi;
    const int cs = e.coff + shifted_cell_of(e.g, px, py, shiftN[2 * r], shiftN[2 * r + 1], e.xmax, e.ymax);
    const int cu = e.coff + cell_of(e.g, px, py);
    cidS[s] = cs;
    cidU[s] = cu;
    int leader, rank, size;
    wave_groups(cs, true, lane, leader, rank, size);
    if (lane == leader) atomicAdd(&cntS[cs], size);
    wave_groups(cu, true, lane, leader, rank, size);
    if (lane == leader) atomicAdd(&cntU[cu], size);
}

/* flow rate after the step: per block of a replica's cells (aligned as in
 * the repair) the sum of the cells' mean x velocity, the state being in
 * cell order (cell c holds slots start[c] .. start[c+1]) */
__global__ void __launch_bounds__(TPB) k_ens_flow_parts(const EnsRep* __restrict__ reps, const int* __restrict__ boff,
                                                        int R, const int* __restrict__ start,
                                                        const double* __restrict__ vx, double* __restrict__ part) {
    __shared__ int sh;
    __shared__ double s0[TPB];
    const int b = blockIdx.x;
    if (threadIdx.x == 0) sh = ens_find(boff, R, b);
    __syncthreads();
    const EnsRep& e = reps[sh];
    const int cl = (b - e.boff) * TPB + (int)threadIdx.x;
    double u = 0.0;
    if (cl < e.g.Nc) {
        const int c = e.coff + cl;
        const int s0i = start[c], s1i = start[c + 1];
        double px = 0.0;
        for (int q = s0i; q < s1i; ++q) px += vx[q];
        if (s1i > s0i) u = px / (double)(s1i - s0i);
    }
    s0[threadIdx.x] = u;
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < (unsigned)s) s0[threadIdx.x] += s0[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[b] = s0[0];
}

/* per replica: Qx = Ly * (mean cell Ux) of this step, accumulated (sum, sum
 * of squares) */
__global__ void __launch_bounds__(TPB) k_ens_flow_acc(const EnsRep* __restrict__ reps, const double* __restrict__ part,
                                                      double* __restrict__ acc) {
    const EnsRep& e = reps[blockIdx.x];
    __shared__ double tot;
    reduce_parts_block(part + e.boff, e.nbc, 1, &tot);
    if (threadIdx.x == 0) {
        const double q = e.g.Ly * tot / (double)e.g.Nc;
        acc[2 * (size_t)blockIdx.x] += q;
        acc[2 * (size_t)blockIdx.x + 1] += q * q;
    }
}

}  // namespace

/* ===================================================================== */
/*                          the ensemble object                          */
/* ===================================================================== */

struct mpcd_gpu_ensemble {
    int R = 0, N = 0, C = 0, B = 0; /* replicas, particles, cells, closure blocks */
    int Z[2] = {0, 0};
    std::vector<mpcd_params> prm;
    std::vector<EnsRep> hreps;
    hipStream_t stream = nullptr;
    /* tables */
    EnsRep* reps = nullptr;
    int *poff = nullptr, *coff = nullptr, *boff = nullptr;
    /* state (two buffer sets: the kick writes the other) */
    double *x = nullptr, *y = nullptr, *vx = nullptr, *vy = nullptr;
    double *x2 = nullptr, *y2 = nullptr, *vx2 = nullptr, *vy2 = nullptr;
    int *pid = nullptr, *pid2 = nullptr, *cidS = nullptr, *cidU = nullptr;
    int parity = 0; /* 0: the state is in x..; 1: the buffers are swapped */
    /* binning: set A (start, key, list) is current when a step begins; the
     * update after the base pass writes set B (startB, keyB, listB) and swaps,
     * the one after the shifted pass swaps back */
    int *cnt = nullptr, *start = nullptr, *cursor = nullptr, *list = nullptr, *tileSum = nullptr;
    unsigned long long* key = nullptr;
    int *startB = nullptr, *listB = nullptr;
    unsigned long long* keyB = nullptr;
    /* the passes' record of moved particles: departures and arrivals per cell,
     * each arrival's rank (-1 for every slot between updates) */
    int *dout = nullptr, *inc = nullptr, *rnk = nullptr;
    int nTiles = 0, narrow = 1, cellCap = 256;
    /* cells */
    double *refPx = nullptr, *refPy = nullptr, *refUx = nullptr, *refUy = nullptr, *pdriveRaw = nullptr;
    double *redPx = nullptr, *redPy = nullptr, *redUx = nullptr, *redUy = nullptr;
    double *duRX = nullptr, *duRY = nullptr, *duEX = nullptr, *duEY = nullptr, *pdrive = nullptr;
    int* mask = nullptr;
    double *part = nullptr, *scal = nullptr;
    /* zones, per layout */
    ZoneRect* rects[2] = {nullptr, nullptr};
    int* zrep[2] = {nullptr, nullptr};
    int* zoneOff[2] = {nullptr, nullptr};
    int* owner[2] = {nullptr, nullptr};
    unsigned char* cover[2] = {nullptr, nullptr};
    int* overCells[2] = {nullptr, nullptr};
    int nOver[2] = {0, 0};
    ZoneIn* ensIn[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}}; /* [parity][layout] */
    int *zsize = nullptr, *zoff = nullptr;
    long long scratchCap = 0;
    int *next = nullptr, *tmpg = nullptr;
    unsigned char *recG = nullptr, *zstate = nullptr;
    double* duG = nullptr;
    ZoneOut* zout = nullptr;
    size_t zsStride = 0;
    int maxC = 0, maxW = 0;
    /* per step */
    long long* stepDev = nullptr;
    long long stepHost = -1; /* the device counter's value after the enqueued work, -1 unknown */
    double* shift = nullptr;
    uint64_t *zseed = nullptr, *seeds = nullptr;
    int* err = nullptr;
    /* flow statistics */
    bool flowOn = false;
    int flowSteps = 0;
    double *flowPart = nullptr, *flowAcc = nullptr;
    /* staging of one replica's interleaved state */
    double* stage = nullptr;
    /* graphs of one step per parity */
    /* graphs of one step per buffer parity, per form: [parity][pre][fuse]
     * (pre: the previous kick streamed this step; fuse: this kick streams
     * the next) */
    hipGraphExec_t graph[2][2][2] = {};
    bool streamedNext = false; /* the state is already streamed for the next step */
    int* cntS = nullptr;       /* the shifted cells' histogram a fused kick counts */

    void check_launch() { HIP_OK(hipGetLastError()); }

    explicit mpcd_gpu_ensemble(const mpcd_params* ps, const uint64_t* keys, int nRep) {
        if (nRep <= 0) throw ApiError(MPCD_E_INVALID, "an ensemble needs at least one replica");
        R = nRep;
        prm.assign(ps, ps + nRep);
        long long n = 0, c = 0, b = 0, z0 = 0, z1 = 0;
        int maxN = 0;
        double maxPpc = 1.0;
        int coverMax = 1;
        int anyLocal = 0;
        std::vector<std::vector<ZoneRect>> lr[2];
        std::vector<std::vector<int32_t>> lown[2];
        hreps.resize((size_t)R);
        for (int r = 0; r < R; ++r) {
            const mpcd_params& p = prm[(size_t)r];
            validate(p);
            if (!p.useIncompressibleRedistribution || !p.useZoneRedistribution || !p.useLiquidClosure)
                throw ApiError(MPCD_E_INVALID, "ensemble replicas run the full step: redistribution, zones and the "
                                               "liquid closure must be on");
            if (p.redistributionEnableSurfaceTopology || p.redistributionWallWettingEnabled ||
                p.useInterfaceVelocityReorientation)
                throw ApiError(MPCD_E_INVALID, "zone pass is restricted to no surface / no wetting / no reorientation");
            if (p.keepMeanFlow) throw ApiError(MPCD_E_INVALID, "keepMeanFlow is not supported in ensembles");
            if (p.n <= 0) throw ApiError(MPCD_E_INVALID, "every replica needs particles");
            EnsRep& e = hreps[(size_t)r];
            e.g = make_geom(p);
            e.b = make_bc_params(p);
            e.cp = make_collide_params(p);
            e.xmax = std::nextafter(p.Lx, 0.0);
            e.ymax = std::nextafter(p.Ly, 0.0);
            e.Vc = (p.Lx / (double)p.Nx) * (p.Ly / (double)p.Ny);
            e.K = p.Kvirial;
            e.rhoT = p.gamma / e.Vc;
            e.gamma = p.gamma;
            e.dtr = p.dt / e.rhoT;
            e.betaEOS = p.betaEOS;
            e.blend = p.smoothPdriveInterzoneBlend;
            e.betaRepair = p.betaRepair;
            e.useOptimal = p.useOptimalBetaRepair;
            e.poff = (int)n;
            e.coff = (int)c;
            e.boff = (int)b;
            e.nbc = (int)nblk(p.Nc);
            e.seedOff = (keys ? keys[r] : 0ULL) * SEED_KEY_MIX;
            n += p.n;
            c += p.Nc;
            b += e.nbc;
            maxN = std::max(maxN, p.n);
            maxPpc = std::max(maxPpc, (double)p.n / (double)p.Nc);
            anyLocal |= make_zone_params(p).useLocal;
            for (int L = 0; L < 2; ++L) {
                std::vector<int32_t> own;
                lr[L].push_back(layout_rects(p, L == 0 ? MPCD_LAYOUT_BASE : MPCD_LAYOUT_SHIFTED, &own));
                lown[L].push_back(std::move(own));
                std::vector<int> cov((size_t)p.Nc, 0);
                for (const auto& q : lr[L].back()) {
                    maxC = std::max(maxC, (q.xe - q.xs) * (q.ye - q.ys));
                    maxW = std::max(maxW, q.xe - q.xs);
                    for (int ix = q.xs; ix < q.xe; ++ix)
                        for (int iy = q.ys; iy < q.ye; ++iy)
                            coverMax = std::max(coverMax, ++cov[(size_t)(iy + p.Ny * ix)]);
                }
                (L == 0 ? z0 : z1) += (long long)lr[L].back().size();
            }
        }
        if (n >= (1LL << 31) || c >= (1LL << 30)) throw ApiError(MPCD_E_INVALID, "ensemble too large");
        N = (int)n;
        C = (int)c;
        B = (int)b;
        Z[0] = (int)z0;
        Z[1] = (int)z1;
        narrow = maxN < (1 << 27) ? 1 : 0;
        /* the fused zone kernel on 16-bit working sets must take every zone */
        const size_t fix16 = zone_fixed_bytes<uint16_t>(maxC, maxW, anyLocal);
        const size_t cellB16 = (size_t)maxC * (4 * sizeof(uint16_t) + 2) + 8 * (size_t)maxW + 16;
        if (!(maxC < 65535 && fix16 + 2048 <= (size_t)ZONE_LDS_FAST && cellB16 + 4096 <= (size_t)ZONE_LDS_FAST))
            throw ApiError(MPCD_E_INVALID, "ensemble zones must fit the fused zone kernel (zone tiles too large)");
        HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        /* tables */
        reps = dalloc<EnsRep>((size_t)R);
        HIP_OK(hipMemcpy(reps, hreps.data(), (size_t)R * sizeof(EnsRep), hipMemcpyHostToDevice));
        std::vector<int> hp((size_t)R + 1), hc((size_t)R + 1), hb((size_t)R + 1);
        for (int r = 0; r < R; ++r) {
            hp[(size_t)r] = hreps[(size_t)r].poff;
            hc[(size_t)r] = hreps[(size_t)r].coff;
            hb[(size_t)r] = hreps[(size_t)r].boff;
        }
        hp[(size_t)R] = N;
        hc[(size_t)R] = C;
        hb[(size_t)R] = B;
        poff = dalloc<int>((size_t)R + 1);
        coff = dalloc<int>((size_t)R + 1);
        boff = dalloc<int>((size_t)R + 1);
        HIP_OK(hipMemcpy(poff, hp.data(), hp.size() * sizeof(int), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(coff, hc.data(), hc.size() * sizeof(int), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(boff, hb.data(), hb.size() * sizeof(int), hipMemcpyHostToDevice));
        /* state and binning */
        for (double** q : {&x, &y, &vx, &vy, &x2, &y2, &vx2, &vy2}) *q = dalloc<double>((size_t)N);
        for (int** q : {&pid, &pid2, &cidS, &cidU, &list, &listB, &rnk}) *q = dalloc<int>((size_t)N);
        key = dalloc<unsigned long long>((size_t)N);
        keyB = dalloc<unsigned long long>((size_t)N);
        for (int** q : {&cnt, &start, &cursor, &startB, &dout, &inc}) *q = dalloc<int>((size_t)C + 1);
        nTiles = (int)((std::max(C, std::max(Z[0], Z[1])) + 1 + SCAN_TILE - 1) / SCAN_TILE);
        tileSum = dalloc<int>((size_t)nTiles);
        cellCap = (int)std::min(4096.0, std::max(256.0, 1.15 * maxPpc * CELL_BLK + 2.0 * CELL_BLK));
        for (double** q : {&refPx, &refPy, &refUx, &refUy, &pdriveRaw, &redPx, &redPy, &redUx, &redUy, &duRX, &duRY,
                           &duEX, &duEY, &pdrive})
            *q = dalloc<double>((size_t)C);
        mask = dalloc<int>((size_t)C);
        part = dalloc<double>(2 * (size_t)B + 8);
        scal = dalloc<double>(4 * (size_t)R + 8);
        flowPart = dalloc<double>((size_t)B + 8);
        flowAcc = dalloc<double>(2 * (size_t)R);
        stage = dalloc<double>(2 * (size_t)maxN);
        stepDev = dalloc<long long>(1);
        shift = dalloc<double>(4 * (size_t)R);  /* this step's, then the next step's */
        zseed = dalloc<uint64_t>(2 * (size_t)R);
        seeds = dalloc<uint64_t>(3 * (size_t)R); /* wall and rotation per replica, then the next wall seeds */
        cntS = dalloc<int>((size_t)C + 1);
        err = dalloc<int>(1);
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, stream, err, 1, 0);
        hipLaunchKernelGGL(k_fill_double, dim3(nblk(2 * (size_t)R)), dim3(TPB), 0, stream, flowAcc, 2 * R, 0.0);
        hipLaunchKernelGGL(k_fill_int, dim3(nblk(N)), dim3(TPB), 0, stream, rnk, N, -1);
        hipLaunchKernelGGL(k_fill_int, dim3(nblk(C + 1)), dim3(TPB), 0, stream, dout, C + 1, 0);
        hipLaunchKernelGGL(k_fill_int, dim3(nblk(C + 1)), dim3(TPB), 0, stream, inc, C + 1, 0);
        check_launch();
        /* interzone masks */
        {
            std::vector<int> hm((size_t)C, 0);
            for (int r = 0; r < R; ++r) {
                const auto m = interzone_mask_host(prm[(size_t)r]);
                hreps[(size_t)r].anyMask = std::count(m.begin(), m.end(), 1) > 0 ? 1 : 0;
                std::copy(m.begin(), m.end(), hm.begin() + hreps[(size_t)r].coff);
            }
            HIP_OK(hipMemcpy(mask, hm.data(), hm.size() * sizeof(int), hipMemcpyHostToDevice));
            HIP_OK(hipMemcpy(reps, hreps.data(), (size_t)R * sizeof(EnsRep), hipMemcpyHostToDevice));
        }
        /* zones of both layouts: rects, replica of each zone, owner and cover
         * maps at each replica's cells, and each replica's zone inputs */
        for (int L = 0; L < 2; ++L) {
            std::vector<ZoneRect> ar;
            std::vector<int> azr, azo((size_t)R), ao((size_t)C), oc;
            std::vector<unsigned char> acv((size_t)C, 0);
            std::vector<int> hasCover((size_t)R, 0);
            for (int r = 0; r < R; ++r) {
                const mpcd_params& p = prm[(size_t)r];
                const int cf = hreps[(size_t)r].coff;
                azo[(size_t)r] = (int)ar.size();
                std::vector<int> cov((size_t)p.Nc, 0);
                for (const auto& q : lr[L][(size_t)r]) {
                    ar.push_back(q);
                    azr.push_back(r);
                    for (int ix = q.xs; ix < q.xe; ++ix)
                        for (int iy = q.ys; iy < q.ye; ++iy) cov[(size_t)(iy + p.Ny * ix)]++;
                }
                for (int cl = 0; cl < p.Nc; ++cl) {
                    ao[(size_t)(cf + cl)] = lown[L][(size_t)r][(size_t)cl];
                    if (cov[(size_t)cl] > 1) {
                        acv[(size_t)(cf + cl)] = 1;
                        oc.push_back(cf + cl);
                        hasCover[(size_t)r] = 1;
                    }
                }
            }
            rects[L] = dalloc<ZoneRect>(ar.size());
            HIP_OK(hipMemcpy(rects[L], ar.data(), ar.size() * sizeof(ZoneRect), hipMemcpyHostToDevice));
            zrep[L] = dalloc<int>(azr.size());
            HIP_OK(hipMemcpy(zrep[L], azr.data(), azr.size() * sizeof(int), hipMemcpyHostToDevice));
            zoneOff[L] = dalloc<int>((size_t)R);
            HIP_OK(hipMemcpy(zoneOff[L], azo.data(), azo.size() * sizeof(int), hipMemcpyHostToDevice));
            owner[L] = dalloc<int>((size_t)C);
            HIP_OK(hipMemcpy(owner[L], ao.data(), ao.size() * sizeof(int), hipMemcpyHostToDevice));
            nOver[L] = (int)oc.size();
            if (!oc.empty()) {
                cover[L] = dalloc<unsigned char>((size_t)C);
                HIP_OK(hipMemcpy(cover[L], acv.data(), acv.size(), hipMemcpyHostToDevice));
                overCells[L] = dalloc<int>(oc.size());
                HIP_OK(hipMemcpy(overCells[L], oc.data(), oc.size() * sizeof(int), hipMemcpyHostToDevice));
            }
            for (int par = 0; par < 2; ++par) {
                std::vector<ZoneIn> zi((size_t)R);
                for (int r = 0; r < R; ++r) {
                    const EnsRep& e = hreps[(size_t)r];
                    ZoneIn& q = zi[(size_t)r];
                    q = ZoneIn{};
                    q.g = e.g;
                    q.zp = make_zone_params(prm[(size_t)r]);
                    q.gcnt = cnt + e.coff;
                    q.gstart = (L == 0 ? start : startB) + e.coff;
                    q.glist = L == 0 ? list : listB;
                    q.gkey = L == 0 ? key : keyB;
                    q.dcid = cidU; /* global cells: the recorded ones get cellOff */
                    q.dout = dout;
                    q.dinc = inc;
                    q.drnk = rnk;
                    q.cellOff = e.coff;
                    q.owner = owner[L] + e.coff;
                    q.gcover = hasCover[(size_t)r] ? cover[L] + e.coff : nullptr;
                    q.x = par ? x2 : x;
                    q.y = par ? y2 : y;
                    q.vx = par ? vx2 : vx;
                    q.vy = par ? vy2 : vy;
                    q.vsx = par ? vx : vx2; /* the side buffer: the other set, free during a pass */
                    q.vsy = par ? vy : vy2;
                }
                ensIn[par][L] = dalloc<ZoneIn>((size_t)R);
                HIP_OK(hipMemcpy(ensIn[par][L], zi.data(), zi.size() * sizeof(ZoneIn), hipMemcpyHostToDevice));
            }
        }
        const int zMax = std::max(Z[0], Z[1]);
        zsize = dalloc<int>((size_t)zMax + 2);
        zoff = dalloc<int>((size_t)zMax + 2);
        scratchCap = (long long)coverMax * (long long)N;
        next = dalloc<int>((size_t)scratchCap + 1);
        tmpg = dalloc<int>((size_t)scratchCap + 1);
        recG = dalloc<unsigned char>((size_t)scratchCap + 1);
        duG = dalloc<double>(2 * (size_t)maxC * (size_t)zMax);
        zsStride = zone_state_bytes<int>(maxC, maxW);
        zstate = dalloc<unsigned char>(zsStride * (size_t)zMax);
        zout = dalloc<ZoneOut>((size_t)zMax);
        HIP_OK(hipStreamSynchronize(stream));
    }

    ~mpcd_gpu_ensemble() {
        for (auto& a : graph)
            for (auto& b : a)
                for (hipGraphExec_t& g : b)
                    if (g) (void)hipGraphExecDestroy(g);
        void* ptrs[] = {reps, poff, coff, boff, x, y, vx, vy, x2, y2, vx2, vy2, pid, pid2, cidS, cidU, cnt, start,
                        cursor, list, tileSum, key, startB, listB, keyB, dout, inc, rnk, refPx, refPy, refUx, refUy, pdriveRaw, redPx, redPy, redUx, redUy,
                        duRX, duRY, duEX, duEY, pdrive, mask, part, scal, rects[0], rects[1], zrep[0], zrep[1],
                        zoneOff[0], zoneOff[1], owner[0], owner[1], cover[0], cover[1], overCells[0], overCells[1],
                        ensIn[0][0], ensIn[0][1], ensIn[1][0], ensIn[1][1], zsize, zoff, next, tmpg, recG, zstate,
                        duG, zout, stepDev, shift, zseed, seeds, err, flowPart, flowAcc, stage, cntS};
        for (void* q : ptrs)
            if (q) (void)hipFree(q);
        if (stream) (void)hipStreamDestroy(stream);
    }

    void drop_graphs() {
        for (auto& a : graph)
            for (auto& b : a)
                for (hipGraphExec_t& g : b)
                    if (g) {
                        HIP_OK(hipGraphExecDestroy(g));
                        g = nullptr;
                    }
    }

    void swap_sets() {
        std::swap(x, x2);
        std::swap(y, y2);
        std::swap(vx, vx2);
        std::swap(vy, vy2);
        std::swap(pid, pid2);
        parity ^= 1;
    }

    /* counting sort of the global cell ids `keyc` (Sim::bin_cells) */
    void bin(const int* keyc, int* counts = nullptr) {
        /* counts: the histogram of keyc, already counted (by a fused kick) */
        int* h = counts ? counts : cnt;
        if (!counts) {
            hipLaunchKernelGGL(k_fill_int, dim3(nblk(C + 1)), dim3(TPB), 0, stream, cnt, C + 1, 0);
            hipLaunchKernelGGL(k_hist, dim3(bin_blocks(N)), dim3(TPB), 0, stream, keyc, N, cnt);
        }
        const int nt = (C + 1 + SCAN_TILE - 1) / SCAN_TILE;
        hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(TPB), 0, stream, h, C + 1, start, tileSum,
                           (const int*)nullptr, cursor);
        hipLaunchKernelGGL(k_scan_add_prefix, dim3(nblk(C + 1)), dim3(TPB), 0, stream, start, C + 1, tileSum);
        hipLaunchKernelGGL(k_scatter, dim3(bin_blocks(N)), dim3(TPB), 0, stream, keyc, pid, N, start, cursor, key);
        hipLaunchKernelGGL(k_segsort, dim3(nblk(C, SEG_CELLS)), dim3(SEG_CELLS), 0, stream, start, C, key, list,
                           narrow, 1);
        check_launch();
    }

    void fields(double* Ux, double* Uy, double* Px, double* Py, double* P) {
        FieldsOut f{};
        f.Ux = Ux;
        f.Uy = Uy;
        f.Px = Px;
        f.Py = Py;
        f.P = P;
        hipLaunchKernelGGL(k_ens_cell_fields, dim3(nblk(C, CELL_BLK)), dim3(CELL_BLK), cellCap * 2 * sizeof(double),
                           stream, reps, coff, R, start, list, C, vx, vy, f, cellCap);
        check_launch();
    }

    /* one snapshot zone pass of layout L over every replica */
    void zone_pass(int L) {
        const int nz = Z[L];
        hipLaunchKernelGGL(k_ens_zone_sizes, dim3(nz + 1), dim3(64), 0, stream, rects[L], zrep[L], reps, nz, start,
                           zsize, err);
        const int nt = (nz + 1 + SCAN_TILE - 1) / SCAN_TILE;
        hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(TPB), 0, stream, zsize, nz + 1, zoff, tileSum,
                           (const int*)nullptr, (int*)nullptr);
        if (nt > 1) {
            hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(TPB), 0, stream, tileSum, nt);
            hipLaunchKernelGGL(k_scan_add, dim3(nblk(nz + 1)), dim3(TPB), 0, stream, zoff, nz + 1, tileSum);
        }
        ZoneArgs a{};
        a.rects = rects[L];
        a.nZones = nz;
        a.zsize = zsize;
        a.zoff = zoff;
        a.scratchCap = scratchCap;
        a.nextG = next;
        a.tmpG = tmpg;
        a.duG = duG;
        a.rec = recG;
        a.zsStride = zsStride;
        a.zstate = zstate;
        a.duStride = 2 * maxC;
        a.pMin = 0;
        a.pMax = 65534;
        a.ldsBytes = ZONE_LDS_FAST;
        a.out = zout;
        a.err = err;
        a.ensRep = zrep[L];
        a.ensIn = ensIn[parity][L];
        a.ensSeed = zseed + (size_t)L * R;
        a.ensZoneOff = zoneOff[L];
        hipLaunchKernelGGL((k_zone_fused<uint16_t, true>), dim3(nz), dim3(FUSED_TPB), ZONE_LDS_FAST, stream, a);
        check_launch();
        if (nOver[L] > 0) {
            hipLaunchKernelGGL(k_overlap_copyback, dim3(nblk(nOver[L])), dim3(TPB), 0, stream, overCells[L], nOver[L],
                               start, list, vx2, vy2, vx, vy);
            check_launch();
        }
    }

    /* the unshifted binning after a zone pass from its recorded moves
     * (Sim::rebin_recorded): the runs are the stayers merged with the
     * arrivals in pid order, what a counting sort of the new cells gives */
    void rebin_recorded() {
        const int nt = (C + 1 + SCAN_TILE - 1) / SCAN_TILE;
        hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(TPB), 0, stream, cnt, C + 1, startB, tileSum, inc,
                           (int*)nullptr, dout);
        hipLaunchKernelGGL(k_scan_add_prefix, dim3(nblk(C + 1)), dim3(TPB), 0, stream, startB, C + 1, tileSum);
        hipLaunchKernelGGL(k_place_arrivals, dim3(nblk((N + PLACE_PER - 1) / PLACE_PER)), dim3(TPB), 0, stream, cidU,
                           pid, rnk, N, startB, cnt, dout, keyB);
        int cpb = REB_CELLS;
        while (cpb > 64 && (double)cpb * (double)N / (double)C * 1.25 > (double)REB_CAP) cpb >>= 1;
        hipLaunchKernelGGL(k_rebin_merge, dim3(nblk(C, cpb)), dim3(REB_CELLS), 0, stream, start, key, cidU, C, cpb,
                           startB, cnt, inc, dout, rnk, keyB, listB, err);
        check_launch();
        std::swap(start, startB);
        std::swap(key, keyB);
        std::swap(list, listB);
    }

    /* one step of every replica, as a fixed launch sequence (the step number
     * is the device counter's); the kick leaves the state in the other
     * buffer set, which the caller swaps in */
    void step_seq(bool pre, bool fuse) {
        hipLaunchKernelGGL(k_ens_seeds, dim3(nblk(R)), dim3(TPB), 0, stream, reps, R, stepDev, shift, zseed, seeds);
        if (!pre) {
            hipLaunchKernelGGL(k_ens_stream, dim3(nblk(N)), dim3(TPB), 0, stream, reps, poff, R, N, seeds, shift, x, y,
                               vx, vy, pid, cidS, cidU, err);
            check_launch();
        }
        /* after a fused stream both histograms are counted (cntS, cnt) */
        bin(cidS, pre ? cntS : nullptr);
        hipLaunchKernelGGL(k_ens_collide, dim3(nblk(C, CELL_BLK)), dim3(CELL_BLK), cellCap * 2 * sizeof(double), stream,
                           reps, coff, R, seeds, start, list, C, vx, vy, cellCap);
        check_launch();
        bin(cidU, pre ? cnt : nullptr); /* the collision moves no particle */
        fields(refUx, refUy, refPx, refPy, pdriveRaw);
        zone_pass(0);
        rebin_recorded();
        zone_pass(1);
        rebin_recorded();
        fields(redUx, redUy, redPx, redPy, nullptr);
        ClosureArgs ca{};
        ca.Nc = C;
        ca.refPx = refPx;
        ca.refPy = refPy;
        ca.refUx = refUx;
        ca.refUy = refUy;
        ca.redPx = redPx;
        ca.redPy = redPy;
        ca.redUx = redUx;
        ca.redUy = redUy;
        ca.duRX = duRX;
        ca.duRY = duRY;
        ca.part = part;
        hipLaunchKernelGGL(k_ens_closure_repair, dim3(B), dim3(TPB), 0, stream, reps, boff, R, ca);
        hipLaunchKernelGGL(k_ens_reduce_beta, dim3(R), dim3(TPB), 0, stream, reps, part, scal);
        hipLaunchKernelGGL(k_ens_pdrive, dim3(nblk(C)), dim3(TPB), 0, stream, reps, coff, R, C, pdriveRaw, mask, pdrive);
        hipLaunchKernelGGL(k_ens_eos, dim3(nblk(C)), dim3(TPB), 0, stream, reps, coff, R, C, pdrive, duEX, duEY);
        if (fuse) {
            hipLaunchKernelGGL(k_fill_int, dim3(nblk(C + 1)), dim3(TPB), 0, stream, cntS, C + 1, 0);
            hipLaunchKernelGGL(k_fill_int, dim3(nblk(C + 1)), dim3(TPB), 0, stream, cnt, C + 1, 0);
            hipLaunchKernelGGL(k_ens_kick_reorder_stream, dim3(nblk(N)), dim3(TPB), 0, stream, reps, poff, R, N,
                               seeds + 2 * (size_t)R, shift + 2 * (size_t)R, list, x, y, vx, vy, pid, x2, y2, vx2, vy2,
                               pid2, duRX, duRY, duEX, duEY, scal, cidS, cidU, cntS, cnt, err);
        } else {
            hipLaunchKernelGGL(k_ens_kick_reorder, dim3(nblk(N)), dim3(TPB), 0, stream, reps, poff, R, N, list, x, y, vx,
                               vy, pid, x2, y2, vx2, vy2, pid2, duRX, duRY, duEX, duEY, scal);
        }
        check_launch();
        if (flowOn && !fuse) {
            hipLaunchKernelGGL(k_ens_flow_parts, dim3(B), dim3(TPB), 0, stream, reps, boff, R, start, vx2, flowPart);
            hipLaunchKernelGGL(k_ens_flow_acc, dim3(R), dim3(TPB), 0, stream, reps, flowPart, flowAcc);
            check_launch();
        }
        hipLaunchKernelGGL(k_ens_step_end, dim3(1), dim3(64), 0, stream, stepDev);
        check_launch();
    }

    /* steps first .. first+k-1 of every replica, enqueued */
    void step(int first, int k, bool useGraph) {
        if (k <= 0) return;
        if (stepHost != (long long)first) {
            hipLaunchKernelGGL(k_ens_set_step, dim3(1), dim3(64), 0, stream, stepDev, (long long)first);
            check_launch();
        }
        for (int j = 0; j < k; ++j) {
            /* every kick but the call's last streams the next step (not
             * while the flow rates are summed: they need the kicked state) */
            const bool pre = streamedNext, fuse = !flowOn && j + 1 < k;
            if (useGraph) {
                hipGraphExec_t& ge = graph[parity][pre][fuse];
                if (!ge) {
                    hipGraph_t gr = nullptr;
                    HIP_OK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
                    step_seq(pre, fuse);
                    HIP_OK(hipStreamEndCapture(stream, &gr));
                    const hipError_t ie = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
                    (void)hipGraphDestroy(gr);
                    HIP_OK(ie);
                }
                HIP_OK(hipGraphLaunch(ge, stream));
            } else {
                step_seq(pre, fuse);
            }
            swap_sets();
            streamedNext = fuse;
            if (flowOn) ++flowSteps;
        }
        stepHost = (long long)first + k;
    }

    void check_err() {
        int h = 0;
        HIP_OK(hipMemcpyAsync(&h, err, sizeof(int), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        if (h) {
            hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, stream, err, 1, 0);
            HIP_OK(hipStreamSynchronize(stream));
            if (h & 1) throw ApiError(MPCD_E_BC_MODE, "a particle hit a wall of unsupported mode");
            if (h & 8) throw ApiError(MPCD_E_CAPACITY, "a zone holds more particles than the 16-bit working set");
            if (h & 4) throw ApiError(MPCD_E_INVALID, "binning update: a cell's stayers do not match its count");
            if (h & 16) throw ApiError(MPCD_E_INVALID, "dense phase: a wait for earlier cells hit its bound");
            throw ApiError(MPCD_E_CAPACITY, "zone scratch exceeded");
        }
    }

    void check_replica(int r) const {
        if (r < 0 || r >= R) throw ApiError(MPCD_E_INVALID, "replica index out of range");
    }

    /* replica r's state from interleaved host arrays, in particle-id order */
    void set_state(int r, const double* hx, const double* hv) {
        check_replica(r);
        const int n = prm[(size_t)r].n, o = hreps[(size_t)r].poff;
        if (hx) {
            HIP_OK(hipMemcpyAsync(stage, hx, 2 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, stream));
            hipLaunchKernelGGL(k_deinterleave, dim3(nblk(n)), dim3(TPB), 0, stream, stage, n, x + o, y + o);
        }
        if (hv) {
            HIP_OK(hipMemcpyAsync(stage + (hx ? 0 : 0), hv, 2 * (size_t)n * sizeof(double), hipMemcpyHostToDevice,
                                  stream));
            hipLaunchKernelGGL(k_deinterleave, dim3(nblk(n)), dim3(TPB), 0, stream, stage, n, vx + o, vy + o);
        }
        hipLaunchKernelGGL(k_iota, dim3(nblk(n)), dim3(TPB), 0, stream, pid + o, n);
        check_launch();
        HIP_OK(hipStreamSynchronize(stream));
    }

    void get_state(int r, double* hx, double* hv) {
        check_replica(r);
        const int n = prm[(size_t)r].n, o = hreps[(size_t)r].poff;
        const double* src[2][2] = {{x + o, y + o}, {vx + o, vy + o}};
        double* dst[2] = {hx, hv};
        for (int k = 0; k < 2; ++k) {
            if (!dst[k]) continue;
            hipLaunchKernelGGL(k_interleave, dim3(nblk(n)), dim3(TPB), 0, stream, src[k][0], src[k][1], pid + o, n,
                               stage);
            check_launch();
            HIP_OK(hipMemcpyAsync(dst[k], stage, 2 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, stream));
            HIP_OK(hipStreamSynchronize(stream));
        }
    }
};

extern "C" {

int mpcd_gpu_sim_set_seed_key(mpcd_gpu_sim* s, uint64_t key) {
    return guarded([&] {
        if (!s) throw ApiError(MPCD_E_INVALID, "null simulation");
        s->seedOff = key * SEED_KEY_MIX;
    });
}

int mpcd_gpu_ensemble_create(const mpcd_params* params, const uint64_t* seedKeys, int nReplicas,
                             mpcd_gpu_ensemble** out) {
    return guarded([&] {
        if (!params || !out) throw ApiError(MPCD_E_INVALID, "null argument");
        *out = nullptr;
        *out = new mpcd_gpu_ensemble(params, seedKeys, nReplicas);
    });
}

int mpcd_gpu_ensemble_destroy(mpcd_gpu_ensemble* e) {
    return guarded([&] { delete e; });
}

int mpcd_gpu_ensemble_info(mpcd_gpu_ensemble* e, int64_t* particles, int32_t* cells, int32_t* zones2) {
    return guarded([&] {
        if (!e) throw ApiError(MPCD_E_INVALID, "null ensemble");
        if (particles) *particles = e->N;
        if (cells) *cells = e->C;
        if (zones2) {
            zones2[0] = e->Z[0];
            zones2[1] = e->Z[1];
        }
    });
}

int mpcd_gpu_ensemble_set_state(mpcd_gpu_ensemble* e, int replica, const double* x, const double* v) {
    return guarded([&] {
        if (!e || !x || !v) throw ApiError(MPCD_E_INVALID, "null argument");
        e->set_state(replica, x, v);
    });
}

int mpcd_gpu_ensemble_get_state(mpcd_gpu_ensemble* e, int replica, double* x, double* v) {
    return guarded([&] {
        if (!e) throw ApiError(MPCD_E_INVALID, "null ensemble");
        e->get_state(replica, x, v);
    });
}

int mpcd_gpu_ensemble_step(mpcd_gpu_ensemble* e, int firstStep, int nSteps, int useGraph) {
    return guarded([&] {
        if (!e) throw ApiError(MPCD_E_INVALID, "null ensemble");
        if (nSteps < 0) throw ApiError(MPCD_E_INVALID, "negative step count");
        e->step(firstStep, nSteps, useGraph != 0);
    });
}

int mpcd_gpu_ensemble_synchronize(mpcd_gpu_ensemble* e) {
    return guarded([&] {
        if (!e) throw ApiError(MPCD_E_INVALID, "null ensemble");
        HIP_OK(hipStreamSynchronize(e->stream));
        e->check_err();
    });
}

int mpcd_gpu_ensemble_flow(mpcd_gpu_ensemble* e, int enable, double* qSum, double* qSumSq, int32_t* nSteps) {
    return guarded([&] {
        if (!e) throw ApiError(MPCD_E_INVALID, "null ensemble");
        if (qSum || qSumSq || nSteps) {
            std::vector<double> h(2 * (size_t)e->R);
            HIP_OK(hipMemcpyAsync(h.data(), e->flowAcc, h.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
            HIP_OK(hipStreamSynchronize(e->stream));
            for (int r = 0; r < e->R; ++r) {
                if (qSum) qSum[r] = h[2 * (size_t)r];
                if (qSumSq) qSumSq[r] = h[2 * (size_t)r + 1];
            }
            if (nSteps) *nSteps = e->flowSteps;
        }
        const bool on = enable != 0;
        if (on != e->flowOn) e->drop_graphs(); /* the launch sequence changes */
        if (on && !e->flowOn) {
            hipLaunchKernelGGL(k_fill_double, dim3(nblk(2 * (size_t)e->R)), dim3(TPB), 0, e->stream, e->flowAcc,
                               2 * e->R, 0.0);
            HIP_OK(hipGetLastError());
            e->flowSteps = 0;
        }
        e->flowOn = on;
    });
}

}  // extern "C"
