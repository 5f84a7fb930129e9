/*
 * biofvm_oracle.c -- CPU oracle for the BioFVM diffusion-decay hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain C, FP64 restatement of the
 * reference algorithm.  It is used by the test-suite as the checker, by the
 * smoke check, and by bench.py as the timed CPU baseline ("port").  The
 * product path (the HIP library under physicell-x_amd/) never links, loads or
 * calls it.
 *
 * What it restates (paths relative to the reference tree):
 *   - solver constants          BioFVM/BioFVM_solvers.cpp:194-217
 *   - Thomas precompute         BioFVM/BioFVM_solvers.cpp:226-329,
 *                               axpy at BioFVM/BioFVM_vector.cpp:331-338
 *   - LOD_3D_BLOCKING step,     BioFVM/BioFVM_solvers.cpp:102-780
 *     world size 1, granularity 1
 *   - LOD_3D_AVX256D step at    BioFVM/BioFVM_solvers.cpp:796-1665
 *     world size 1 (no Dirichlet pass after the z sweep)
 *   - Dirichlet faces           BioFVM/BioFVM_microenvironment.cpp:348-446
 *   - Dirichlet value table     BioFVM/BioFVM_microenvironment.cpp:218-226
 *     (add_dirichlet_node),     :632 (default 1.0), :2649-2786 (face order)
 *   - agent constants           BioFVM/BioFVM_basic_agent.cpp:224-274
 *   - secretion / uptake        BioFVM/BioFVM_basic_agent.cpp:509-559,
 *     in serial agent order     BioFVM/BioFVM_microenvironment.cpp:1370-1384
 *   - gradient field            BioFVM/BioFVM_microenvironment.cpp:1454-1615
 *     (compute_all_gradient_vectors)
 *   - bulk sources / sinks      BioFVM/BioFVM_microenvironment.cpp:1327-1368
 *   - mesh node count           BioFVM/BioFVM_mesh.cpp:1471-1473
 *   - nearest voxel             BioFVM/BioFVM_mesh.cpp:1825-1843,
 *                               BioFVM/BioFVM_basic_agent.cpp:177-186,
 *                               BioFVM/BioFVM_mesh.cpp:242-251
 *
 * Layout is the reference's: idx(i,j,k,d) = ((i*ny + j)*nz + k)*ns + d
 * (BioFVM/BioFVM_mesh.cpp:1157), x slowest, substrate fastest.
 *
 * Floating point.  The reference is built with g++ -O3 -march=native, which
 * contracts every `a += b*c` / `a -= b*c` statement into one fused
 * multiply-add.  This file is compiled with -ffp-contract=off and writes those
 * contractions out as fma() calls, so every rounding step is explicit.  All
 * divisions are true IEEE divisions, as in the reference.
 *
 * Pinning: each function is checked by tests/test_oracle_*.py against
 * exact-rational fixtures (tests/golden/), dense linear solves, Neumann
 * cosine eigenmodes, the D=0 and lambda=0 closed forms, the backward-Euler
 * residual of the agent update and the internalised-mass invariant.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(i, j, k, d) ((((size_t)(i) * ny + (j)) * nz + (k)) * ns + (d))

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* Cartesian_Mesh::resize node count: (int) ceil(eps + (end-start)/dx), eps=1e-16
 * (BioFVM/BioFVM_mesh.cpp:1469-1473). */
int orc_mesh_nodes(double start, double end, double dx)
{
    return (int)ceil(1e-16 + (end - start) / dx);
}

/* Solver constants, BioFVM/BioFVM_solvers.cpp:194-217, one entry per
 * substrate: c1 = D*dt/dx/dx, c2 = lambda*dt/3, c3 = 1+c1+c1+c2,
 * c3a = 1+c1+c2, each built in the reference's statement order. */
void orc_thomas_constants(int ns, const double *D, const double *lambda,
                          double dt, double dx, double *c1, double *c2,
                          double *c3, double *c3a)
{
    for (int d = 0; d < ns; d++) {
        double a = D[d];
        a *= dt;
        a /= dx;
        a /= dx;
        double b = lambda[d];
        b *= dt;
        b /= 3.0;
        double c = 1.0;
        c += a;
        c += a;
        c += b;
        double ca = 1.0;
        ca += a;
        ca += b;
        c1[d] = a;
        c2[d] = b;
        c3[d] = c;
        c3a[d] = ca;
    }
}

/* Thomas coefficients for one axis of n nodes (BioFVM/BioFVM_solvers.cpp:
 * 226-329).  Output arrays are [n][ns] like the reference's
 * vector<vector<double>>.  denom is the LU pivot, cp the c' coefficient. */
void orc_thomas_axis(int n, int ns, const double *c1, const double *c2,
                     const double *c3, const double *c3a, double *denom,
                     double *cp)
{
    for (int i = 0; i < n; i++)
        for (int d = 0; d < ns; d++) {
            cp[(size_t)i * ns + d] = c1[d] * -1.0;
            denom[(size_t)i * ns + d] = c3[d];
        }
    for (int d = 0; d < ns; d++) {
        denom[d] = c3a[d];
        denom[(size_t)(n - 1) * ns + d] = c3a[d];
        if (n == 1) {
            denom[d] = 1.0;
            denom[d] += c2[d];
        }
        cp[d] /= denom[d];
    }
    for (int i = 1; i < n; i++)
        for (int d = 0; d < ns; d++) {
            size_t o = (size_t)i * ns + d;
            denom[o] = fma(c1[d], cp[o - ns], denom[o]);
            cp[o] /= denom[o];
        }
}

/* Dirichlet value planes.  The reference keeps one value vector per voxel,
 * initialised to 1.0 (BioFVM/BioFVM_microenvironment.cpp:632) and overwritten
 * by add_dirichlet_node (:218-226) for every voxel of every registered face,
 * in the order xmin, xmax, ymin, ymax, zmin, zmax (:2649-2786).  Only
 * boundary voxels are ever read, so the table is stored as six face planes:
 *   plane 0/1 (x = 0 / nx-1): [j][k][d], ny*nz*ns values
 *   plane 2/3 (y = 0 / ny-1): [i][k][d], nx*nz*ns values
 *   plane 4/5 (z = 0 / nz-1): [i][j][d], nx*ny*ns values
 * A voxel lying on several faces has one entry in each of their planes; the
 * replay below writes all of them, so they always agree. */
static void plane_write(double *planes[6], int nx, int ny, int nz, int ns,
                        int i, int j, int k, const double *val)
{
    for (int d = 0; d < ns; d++) {
        if (i == 0) planes[0][((size_t)j * nz + k) * ns + d] = val[d];
        if (i == nx - 1) planes[1][((size_t)j * nz + k) * ns + d] = val[d];
        if (j == 0) planes[2][((size_t)i * nz + k) * ns + d] = val[d];
        if (j == ny - 1) planes[3][((size_t)i * nz + k) * ns + d] = val[d];
        if (k == 0) planes[4][((size_t)i * ny + j) * ns + d] = val[d];
        if (k == nz - 1) planes[5][((size_t)i * ny + j) * ns + d] = val[d];
    }
}

void orc_dirichlet_planes(int nx, int ny, int nz, int ns,
                          const unsigned char *face_registered,
                          const double *face_values, double *p0, double *p1,
                          double *p2, double *p3, double *p4, double *p5)
{
    double *planes[6] = {p0, p1, p2, p3, p4, p5};
    size_t sz[6] = {(size_t)ny * nz * ns, (size_t)ny * nz * ns,
                    (size_t)nx * nz * ns, (size_t)nx * nz * ns,
                    (size_t)nx * ny * ns, (size_t)nx * ny * ns};
    for (int f = 0; f < 6; f++)
        for (size_t q = 0; q < sz[f]; q++) planes[f][q] = 1.0;
    /* xmin, xmax: k outer, j inner (BioFVM_microenvironment.cpp:2663-2700) */
    for (int f = 0; f < 2; f++) {
        if (!face_registered[f]) continue;
        int I = f == 0 ? 0 : nx - 1;
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++)
                plane_write(planes, nx, ny, nz, ns, I, j, k,
                            face_values + (size_t)f * ns);
    }
    /* ymin, ymax: k outer, i inner (:2703-2740) */
    for (int f = 2; f < 4; f++) {
        if (!face_registered[f]) continue;
        int J = f == 2 ? 0 : ny - 1;
        for (int k = 0; k < nz; k++)
            for (int i = 0; i < nx; i++)
                plane_write(planes, nx, ny, nz, ns, i, J, k,
                            face_values + (size_t)f * ns);
    }
    /* zmin, zmax: j outer, i inner (:2743-2784) */
    for (int f = 4; f < 6; f++) {
        if (!face_registered[f]) continue;
        int K = f == 4 ? 0 : nz - 1;
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++)
                plane_write(planes, nx, ny, nz, ns, i, j, K,
                            face_values + (size_t)f * ns);
    }
}

/* apply_dirichlet_boundaries_conditions(rank=0, size=1)
 * (BioFVM/BioFVM_microenvironment.cpp:348-446).  Its "#pragma parallel for"
 * lines are not OpenMP pragmas, so the loops run serially here as well. */
void orc_apply_dirichlet(double *u, int nx, int ny, int nz, int ns,
                         const unsigned char *active, const double *p0,
                         const double *p1, const double *p2, const double *p3,
                         const double *p4, const double *p5)
{
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++)
            for (int d = 0; d < ns; d++)
                if (active[d])
                    u[IDX(0, j, k, d)] = p0[((size_t)j * nz + k) * ns + d];
    for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++)
            for (int d = 0; d < ns; d++)
                if (active[d])
                    u[IDX(nx - 1, j, k, d)] = p1[((size_t)j * nz + k) * ns + d];
    for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz; k++) {
            for (int d = 0; d < ns; d++)
                if (active[d])
                    u[IDX(i, 0, k, d)] = p2[((size_t)i * nz + k) * ns + d];
            for (int d = 0; d < ns; d++)
                if (active[d])
                    u[IDX(i, ny - 1, k, d)] = p3[((size_t)i * nz + k) * ns + d];
        }
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            for (int d = 0; d < ns; d++)
                if (active[d])
                    u[IDX(i, j, 0, d)] = p4[((size_t)i * ny + j) * ns + d];
            for (int d = 0; d < ns; d++)
                if (active[d])
                    u[IDX(i, j, nz - 1, d)] = p5[((size_t)i * ny + j) * ns + d];
        }
}

/* One LOD_3D_BLOCKING step at world size 1 (BioFVM/BioFVM_solvers.cpp:
 * 339-780): Dirichlet, x sweep (forward :352-382, back :543-562), Dirichlet,
 * y sweep (:684-729), Dirichlet, z sweep (:733-777), Dirichlet.  The OpenMP
 * work split is the reference's: x lines over (j,k), y over collapse(i,k),
 * z over collapse(i,j). */
void orc_lod_step_variant(double *u, int nx, int ny, int nz, int ns, const double *c1,
                          const double *denx, const double *cpx, const double *deny,
                          const double *cpy, const double *denz, const double *cpz,
                          const unsigned char *active, const double *p0,
                          const double *p1, const double *p2, const double *p3,
                          const double *p4, const double *p5, int final_dirichlet);

void orc_lod_step(double *u, int nx, int ny, int nz, int ns, const double *c1,
                  const double *denx, const double *cpx, const double *deny,
                  const double *cpy, const double *denz, const double *cpz,
                  const unsigned char *active, const double *p0,
                  const double *p1, const double *p2, const double *p3,
                  const double *p4, const double *p5)
{
    orc_lod_step_variant(u, nx, ny, nz, ns, c1, denx, cpx, deny, cpy, denz, cpz, active,
                         p0, p1, p2, p3, p4, p5, 1);
}

/* The step with or without the Dirichlet pass after the z sweep.
 * final_dirichlet = 1 is LOD_3D_BLOCKING (BioFVM/BioFVM_solvers.cpp:778);
 * final_dirichlet = 0 is LOD_3D_AVX256D at world size 1, whose three passes
 * come before the x, y and z sweeps only (:1072, :1510, :1618) and whose
 * recurrences at rank 0 are the same statements. */
void orc_lod_step_variant(double *u, int nx, int ny, int nz, int ns, const double *c1,
                          const double *denx, const double *cpx, const double *deny,
                          const double *cpy, const double *denz, const double *cpz,
                          const unsigned char *active, const double *p0,
                          const double *p1, const double *p2, const double *p3,
                          const double *p4, const double *p5, int final_dirichlet)
{
    const size_t ij = (size_t)ny * nz * ns; /* thomas_i_jump */
    const size_t jj = (size_t)nz * ns;      /* thomas_j_jump */
    const size_t kj = (size_t)ns;           /* thomas_k_jump */
    const long nlines_x = (long)ny * nz;

    orc_apply_dirichlet(u, nx, ny, nz, ns, active, p0, p1, p2, p3, p4, p5);

    /* x forward elimination */
#pragma omp parallel for
    for (long line = 0; line < nlines_x; line++) {
        size_t index = (size_t)line * ns;
        size_t dec = index;
        for (int d = 0; d < ns; d++) u[index + d] /= denx[d];
        for (int i = 1; i < nx; i++) {
            size_t inc = dec + ij;
            for (int d = 0; d < ns; d++)
                u[inc + d] = fma(c1[d], u[dec + d], u[inc + d]);
            for (int d = 0; d < ns; d++)
                u[inc + d] /= denx[(size_t)i * ns + d];
            dec = inc;
        }
    }
    /* x back substitution */
#pragma omp parallel for
    for (long line = 0; line < nlines_x; line++) {
        size_t aux = (size_t)(nx - 1) * ij + (size_t)line * ns;
        for (int i = nx - 2; i >= 0; i--) {
            size_t dec = aux - ij;
            for (int d = 0; d < ns; d++)
                u[dec + d] = fma(-cpx[(size_t)i * ns + d], u[aux + d], u[dec + d]);
            aux = dec;
        }
    }

    orc_apply_dirichlet(u, nx, ny, nz, ns, active, p0, p1, p2, p3, p4, p5);

    /* y sweep */
#pragma omp parallel for collapse(2)
    for (int i = 0; i < nx; i++)
        for (int k = 0; k < nz; k++) {
            size_t index = (size_t)i * ij + (size_t)k * kj;
            for (int d = 0; d < ns; d++) u[index + d] /= deny[d];
            for (int j = 1; j < ny; j++) {
                size_t inc = index + jj;
                for (int d = 0; d < ns; d++)
                    u[inc + d] = fma(c1[d], u[index + d], u[inc + d]);
                for (int d = 0; d < ns; d++)
                    u[inc + d] /= deny[(size_t)j * ns + d];
                index = inc;
            }
            index = (size_t)i * ij + (size_t)k * kj + jj * (size_t)(ny - 1);
            for (int j = ny - 2; j >= 0; j--) {
                size_t dec = index - jj;
                for (int d = 0; d < ns; d++)
                    u[dec + d] = fma(-cpy[(size_t)j * ns + d], u[index + d], u[dec + d]);
                index = dec;
            }
        }

    orc_apply_dirichlet(u, nx, ny, nz, ns, active, p0, p1, p2, p3, p4, p5);

    /* z sweep */
#pragma omp parallel for collapse(2)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            size_t index = (size_t)i * ij + (size_t)j * jj;
            for (int d = 0; d < ns; d++) u[index + d] /= denz[d];
            for (int k = 1; k < nz; k++) {
                size_t inc = index + kj;
                for (int d = 0; d < ns; d++)
                    u[inc + d] = fma(c1[d], u[index + d], u[inc + d]);
                for (int d = 0; d < ns; d++)
                    u[inc + d] /= denz[(size_t)k * ns + d];
                index = inc;
            }
            index = (size_t)i * ij + (size_t)j * jj + kj * (size_t)(nz - 1);
            for (int k = nz - 2; k >= 0; k--) {
                size_t dec = index - kj;
                for (int d = 0; d < ns; d++)
                    u[dec + d] = fma(-cpz[(size_t)k * ns + d], u[index + d], u[dec + d]);
                index = dec;
            }
        }

    if (final_dirichlet)
        orc_apply_dirichlet(u, nx, ny, nz, ns, active, p0, p1, p2, p3, p4, p5);
}

/* Basic_Agent::set_internal_uptake_constants(dt)
 * (BioFVM/BioFVM_basic_agent.cpp:224-274) for n agents.  Per-agent inputs:
 * volume V[n]; per agent and substrate [n][ns]: secretion S, uptake U,
 * net export E, saturation density rho_star.  W is the voxel volume dV.
 * Outputs [n][ns]: c1 (temp1), c2 (temp2), e1 (temp_export1), e2
 * (temp_export2). */
void orc_agent_constants(int n, int ns, double dt, double W, const double *V,
                         const double *S, const double *U, const double *E,
                         const double *rho_star, double *c1, double *c2,
                         double *e1, double *e2)
{
    for (int a = 0; a < n; a++) {
        double f = dt * V[a] / W;
        for (int d = 0; d < ns; d++) {
            size_t o = (size_t)a * ns + d;
            double t1 = 0.0;
            t1 += S[o];
            t1 *= rho_star[o];
            t1 *= f;
            double t2 = 1.0;
            t2 = fma(f, S[o], t2);
            t2 = fma(f, U[o], t2);
            double x1 = E[o];
            x1 *= dt;
            double x2 = x1;
            x2 /= W;
            c1[o] = t1;
            c2[o] = t2;
            e1[o] = x1;
            e2[o] = x2;
        }
    }
}

/* Microenvironment::simulate_cell_sources_and_sinks(dt) run serially in
 * ascending agent index (BioFVM/BioFVM_microenvironment.cpp:1370-1384), each
 * agent doing Basic_Agent::simulate_secretion_and_uptake
 * (BioFVM/BioFVM_basic_agent.cpp:509-559).  voxel[] holds the reference
 * voxel index (current_voxel_index).  When internalized is not NULL the
 * per-agent internalised totals [n][ns] are tracked as with
 * track_internalized_substrates_in_each_agent (:520-532, :553-556). */
void orc_secretion_uptake(double *u, int ns, int n, const int *voxel,
                          const unsigned char *active, const double *c1,
                          const double *c2, const double *e1, const double *e2,
                          double W, double *internalized)
{
    for (int a = 0; a < n; a++) {
        if (!active[a]) continue;
        double *rho = u + (size_t)voxel[a] * ns;
        const size_t o = (size_t)a * ns;
        if (internalized) {
            for (int d = 0; d < ns; d++) {
                double t = 1.0;
                t -= c2[o + d];
                t *= rho[d];
                t += c1[o + d];
                t /= c2[o + d];
                t *= W;
                internalized[o + d] -= t;
            }
        }
        for (int d = 0; d < ns; d++) rho[d] += c1[o + d];
        for (int d = 0; d < ns; d++) rho[d] /= c2[o + d];
        for (int d = 0; d < ns; d++) rho[d] += e2[o + d];
        if (internalized)
            for (int d = 0; d < ns; d++) internalized[o + d] -= e1[o + d];
    }
}

/* Total substrate per species, sum over voxels of rho * dV, accumulated with
 * Neumaier compensation in voxel order.  The reference has no such function
 * on this path; this is the oracle's definition of "total mass". */
void orc_total_mass(const double *u, size_t nvox, int ns, double dV, double *out)
{
    for (int d = 0; d < ns; d++) {
        double s = 0.0, c = 0.0;
        for (size_t v = 0; v < nvox; v++) {
            double x = u[v * ns + d];
            double t = s + x;
            if (fabs(s) >= fabs(x))
                c += (s - t) + x;
            else
                c += (x - t) + s;
            s = t;
        }
        out[d] = (s + c) * dV;
    }
}

/* Basic_Agent::update_voxel_index (BioFVM/BioFVM_basic_agent.cpp:177-186):
 * -1 when the position lies outside the bounding box
 * (General_Mesh::is_position_valid, BioFVM/BioFVM_mesh.cpp:242-251), else
 * Cartesian_Mesh::nearest_voxel_index (BioFVM/BioFVM_mesh.cpp:1825-1843) with
 * world size 1 (local_bounding_box[0] == bounding_box[0]).  bbox is
 * {xmin, ymin, zmin, xmax, ymax, zmax}. */
void orc_voxel_indices(int n, const double *pos, const double *bbox,
                       double dx, double dy, double dz, int nx, int ny, int nz,
                       int *out)
{
    for (int a = 0; a < n; a++) {
        double x = pos[3 * a], y = pos[3 * a + 1], z = pos[3 * a + 2];
        if (x < bbox[0] || x > bbox[3] || y < bbox[1] || y > bbox[4] ||
            z < bbox[2] || z > bbox[5]) {
            out[a] = -1;
            continue;
        }
        unsigned int i = (unsigned int)floor((x - bbox[0]) / dx);
        unsigned int j = (unsigned int)floor((y - bbox[1]) / dy);
        unsigned int k = (unsigned int)floor((z - bbox[2]) / dz);
        if (i >= (unsigned)nx) i = nx - 1;
        if (j >= (unsigned)ny) j = ny - 1;
        if (k >= (unsigned)nz) k = nz - 1;
        out[a] = (int)((i * (unsigned)ny + j) * (unsigned)nz + k);
    }
}


/* Gradient field, compute_all_gradient_vectors
 * (BioFVM/BioFVM_microenvironment.cpp:1454-1615): per voxel and substrate the
 * central difference (u[+1] - u[-1]) / (2 h) inside and the one-sided
 * differences (u[1] - u[0]) / h, (u[n-1] - u[n-2]) / h at the two ends, along
 * x, y and z, each as one subtraction followed by one division.
 * g[(n*ns + q)*3 + c].  An axis with a single node gets 0: the reference
 * skips z when nz == 1 (:1560-1562) and reads outside the field for x or y. */
void orc_gradients(int nx, int ny, int nz, int ns, double dx, double dy, double dz,
                   const double *u, double *g)
{
    const long long jump[3] = {(long long)ny * nz * ns, (long long)nz * ns, (long long)ns};
    const int len[3] = {nx, ny, nz};
    const double h[3] = {dx, dy, dz};
    const double h2[3] = {2.0 * dx, 2.0 * dy, 2.0 * dz};
    #pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                const long long n = ((long long)i * ny + j) * nz + k;
                const int idx[3] = {i, j, k};
                for (int q = 0; q < ns; ++q) {
                    const long long a = n * ns + q;
                    for (int c = 0; c < 3; ++c) {
                        double v;
                        if (len[c] == 1)
                            v = 0.0;
                        else if (idx[c] == 0)
                            v = (u[a + jump[c]] - u[a]) / h[c];
                        else if (idx[c] == len[c] - 1)
                            v = (u[a] - u[a - jump[c]]) / h[c];
                        else
                            v = (u[a + jump[c]] - u[a - jump[c]]) / h2[c];
                        g[a * 3 + c] = v;
                    }
                }
            }
}

/* Bulk sources and sinks, simulate_bulk_sources_and_sinks
 * (BioFVM/BioFVM_microenvironment.cpp:1327-1368), with the per-voxel rates
 * already evaluated: S = supply rate, T = supply target, U = uptake rate, all
 * [nvox][ns].  Per value: t2 = T*S; rho += dt*t2 (one statement, contracted
 * by the reference build: fma); t3 = U + S; t3 *= dt; t3 += 1; rho /= t3
 * (separate statements, separate roundings). */
void orc_bulk_sources(double *u, size_t nvox, int ns, double dt, const double *S,
                      const double *T, const double *U)
{
    const long long total = (long long)nvox * ns;
    #pragma omp parallel for schedule(static)
    for (long long a = 0; a < total; ++a) {
        const double t2 = T[a] * S[a];
        double rho = fma(dt, t2, u[a]);
        double t3 = U[a] + S[a];
        t3 = t3 * dt;
        t3 = t3 + 1.0;
        u[a] = rho / t3;
    }
}
