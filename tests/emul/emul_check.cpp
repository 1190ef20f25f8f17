// Runs the product C-ABI, built for the CPU against tests/emul/hip/hip_runtime.h
// with AddressSanitizer, over every flavour, precision, x-pass shape variant,
// graph path, tangent-linear/adjoint, Newton and recurrence-search entry point at small sizes,
// and checks the results against the FP64 oracle.  Exit status 0 = all passed.
// TEST INFRASTRUCTURE ONLY.
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <complex>
#include <vector>
#include "oracle.h"
#include "psg.h"
#include "psg_adjoint.h"
#include "psg_newton.h"
#include "psg_rec.h"
#include "psg_stats.h"

typedef std::vector<double> V;
static int g_fail = 0, g_pass = 0;

static void check(bool ok, const char *what, double v) {
  printf("%s %-58s %.3e\n", ok ? "PASS" : "FAIL", what, v);
  fflush(stdout);
  if (ok) g_pass++;
  else g_fail++;
}
static double rel(const V &a, const V &b) {
  double m = 0, n = 0;
  for (size_t i = 0; i < a.size(); i++) { m = fmax(m, fabs(a[i] - b[i])); n = fmax(n, fabs(b[i])); }
  return m / (n > 0 ? n : 1);
}
static psg_config cfg3(int flav, int nx, int ny, int nz, double alpha, int prec, int batch, double v2, double dt,
                       int force) {
  psg_config c;
  memset(&c, 0, sizeof c);
  c.flavour = flav; c.precision = prec; c.batch = batch; c.nx = nx; c.ny = ny; c.nz = nz; c.alpha = alpha;
  c.v2 = v2; c.dt = dt; c.force = force; c.kf = 4.0;
  c.ampfor = force == PSG_FORCE_DNS ? -4.0 : (force == PSG_FORCE_GMRES ? -1.0 : (force == PSG_FORCE_KOLMO ? -2.0 : 0));
  c.ri = 1.0; c.sc = 1.0; c.theta = M_PI / 2;
  return c;
}
static V ic3(int nx, int ny, int nz, double alpha, uint64_t seed, int nf) {
  or_grid g;
  or_grid3(&g, nx, ny, nz, alpha);
  V s(2 * nf * g.nkt, 0.0);
  orc *p = (orc *)s.data();
  or_ic_solenoidal(&g, seed, 0.5, 0.1, p, p + g.nkt, p + 2 * g.nkt, nf == 4 ? p + 3 * g.nkt : nullptr);
  return s;
}

static void ns3d_case(const char *name, int flav, int nx, int ny, int nz, double alpha, int prec, int variant,
                      int graph, int steps, int force, double tol, int zvariant = 0, int slab = 0, int batch = 1) {
  psg_config c = cfg3(flav, nx, ny, nz, alpha, prec, batch, flav == PSG_EULER3D ? 0.0 : 0.02, 0.004, force);
  psg_ctx *ctx;
  if (psg_create(&c, &ctx)) { check(false, name, 1); return; }
  psg_set_variant(ctx, PSG_PASS_XPASS, variant);
  psg_set_variant(ctx, PSG_PASS_ZFWD, zvariant);
  psg_set_variant(ctx, PSG_PASS_SLAB, slab);
  psg_set_graph(ctx, graph);
  std::vector<V> st, got;
  for (int m = 0; m < batch; m++) {
    st.push_back(ic3(nx, ny, nz, alpha, 11 + m, 3));
    got.push_back(V(st[m].size()));
    psg_set_state(ctx, m, st[m].data());
  }
  psg_run(ctx, steps);
  for (int m = 0; m < batch; m++) psg_get_state(ctx, m, got[m].data());
  double e, z, sp[64];
  psg_diag(ctx, 0, &e, &z, sp);
  psg_destroy(ctx);
  or3_ctx *o = or3_create(nx, ny, nz, alpha, flav == PSG_EULER3D ? 1 : 0, c.v2, c.dt, force, c.kf, c.ampfor);
  double worst = 0.0;
  for (int m = 0; m < batch; m++) {
    const long nkt = (long)st[m].size() / 6;
    V want = st[m];
    orc *w = (orc *)want.data();
    or3_run(o, w, w + nkt, w + 2 * nkt, steps);
    worst = std::max(worst, rel(got[m], want));
  }
  or3_destroy(o);
  check(worst < tol, name, worst);
}

static void bous_case(const char *name, int nx, int ny, int nz, double alpha, int prec, int batch, int graph,
                      int variant = 0) {
  psg_config c = cfg3(PSG_BOUS3D, nx, ny, nz, alpha, prec, batch, 0.01, 0.004, PSG_FORCE_KOLMO);
  psg_ctx *ctx;
  if (psg_create(&c, &ctx)) { check(false, name, 1); return; }
  psg_set_graph(ctx, graph);
  psg_set_variant(ctx, PSG_PASS_XPASS, variant);
  std::vector<V> st;
  for (int m = 0; m < batch; m++) { st.push_back(ic3(nx, ny, nz, alpha, 40 + m, 4)); psg_set_state(ctx, m, st[m].data()); }
  psg_run(ctx, 4);
  const long nkt = (long)st[0].size() / 8;
  orb_ctx *o = orb_create(nx, ny, nz, alpha, c.v2, c.sc, c.ri, c.theta, c.dt, c.kf, c.ampfor);
  double worst = 0;
  for (int m = 0; m < batch; m++) {
    V got(st[m].size()), want = st[m];
    psg_get_state(ctx, m, got.data());
    orc *w = (orc *)want.data();
    orb_run(o, w, w + nkt, w + 2 * nkt, w + 3 * nkt, 4);
    worst = fmax(worst, rel(got, want));
  }
  // dz/dt and physical fields of member 0 through the same context
  V d(st[0].size()), dw(st[0].size());
  psg_timederiv(ctx, 0, st[0].data(), d.data());
  orc *s0 = (orc *)st[0].data(), *o0 = (orc *)dw.data();
  orb_rhs(o, s0, s0 + nkt, s0 + 2 * nkt, s0 + 3 * nkt, o0, o0 + nkt, o0 + 2 * nkt, o0 + 3 * nkt);
  V phys((size_t)7 * nx * ny * nz);
  psg_physical(ctx, 0, phys.data());
  orb_destroy(o);
  psg_destroy(ctx);
  check(worst < (prec ? 1e-4 : 1e-11), name, worst);
  check(rel(d, dw) < (prec ? 1e-4 : 1e-12), "  bous timederiv", rel(d, dw));
}

static void ns2d_case(const char *name, int n, int prec, int batch, int graph, int steps) {
  psg_config c = cfg3(PSG_NS2D, n, n, 1, 1.0, prec, batch, 1 / 40.0, 0.005, PSG_FORCE_KOLMO);
  psg_ctx *ctx;
  if (psg_create(&c, &ctx)) { check(false, name, 1); return; }
  psg_set_graph(ctx, graph);
  or_grid g;
  or_grid2(&g, n, n, 1.0);
  std::vector<V> st;
  for (int m = 0; m < batch; m++) {
    st.push_back(V(2 * g.nkt));
    or_ic_ns2d_cfg(&g, 5 + m, 1.0, (orc *)st[m].data());
    psg_set_state(ctx, m, st[m].data());
  }
  psg_run(ctx, steps);
  or2_ctx *o = or2_create(n, n, 1.0, c.v2, c.dt, 4, -2.0);
  double worst = 0;
  for (int m = 0; m < batch; m++) {
    V got(st[m].size()), want = st[m];
    psg_get_state(ctx, m, got.data());
    or2_run(o, (orc *)want.data(), steps);
    worst = fmax(worst, rel(got, want));
  }
  V nl(st[0].size()), nlw(st[0].size()), phys((size_t)3 * n * n);
  psg_nonlin(ctx, 0, st[0].data(), nl.data());
  or2_nonlin(o, (orc *)st[0].data(), (orc *)nlw.data());
  psg_physical(ctx, 0, phys.data());
  or2_destroy(o);
  psg_destroy(ctx);
  check(worst < (prec ? 1e-4 : 1e-11), name, worst);
  check(rel(nl, nlw) < (prec ? 1e-4 : 1e-12), "  2d nonlin", rel(nl, nlw));
}

static double ip(const V &a, const V &b, const std::vector<int> &L, long nkt) {
  double s = 0;
  for (size_t i = 0; i < a.size(); i += 2) {
    const long ik = (long)(i / 2) % nkt;
    if (L[ik]) s += a[i] * b[i] + a[i + 1] * b[i + 1];
  }
  return s;
}

static void tladj_case(int n, int prec) {
  psg_config c = cfg3(PSG_NS3D, n, n, n, 1.0, prec, 1, 0.02, 0.004, PSG_FORCE_KOLMO);
  psg_ctx *ctx;
  if (psg_create(&c, &ctx)) { check(false, "tl/adjoint create", 1); return; }
  or_grid g;
  or_grid3(&g, n, n, n, 1.0);
  std::vector<int> L(g.nkt);
  or_tables(&g, nullptr, nullptr, nullptr, L.data());
  V z = ic3(n, n, n, 1.0, 3, 3), v = ic3(n, n, n, 1.0, 4, 3), w = ic3(n, n, n, 1.0, 5, 3);
  V zT(z.size()), jv(z.size()), jtw(z.size());
  const int nst = 7;
  psg_tangent_steps(ctx, z.data(), v.data(), nst, zT.data(), jv.data());
  psg_adjoint(ctx, z.data(), w.data(), nst, 3, nullptr, jtw.data());
  const double a = ip(jv, w, L, g.nkt), b = ip(v, jtw, L, g.nkt);
  check(fabs(a - b) <= (prec ? 1e-4 : 1e-11) * fabs(a), prec ? "tl/adjoint dot product fp32" : "tl/adjoint dot product fp64",
        fabs(a - b) / fabs(a));
  double J;
  V grad(z.size());
  psg_residual_grad(ctx, z.data(), nst, 2, 0.1, 0.0, 0.0, &J, grad.data(), nullptr);
  check(J > 0 && std::isfinite(J), "  residual grad finite", J);
  // 3D Newton-tangent route and one Newton iteration with TL Jv (bounds only)
  V z1(z.size()), dz1(z.size());
  psg_tangent(ctx, z.data(), v.data(), 5 * c.dt, c.dt, PSG_TIME_LASTSTEP, z1.data(), dz1.data());
  check(rel(dz1, dz1) == 0 && std::isfinite(dz1[0]), "  psg_tangent 3d", 0);
  psg_destroy(ctx);
}

static void newton2d_case(int jv) {
  const int n = 32;
  psg_config c = cfg3(PSG_NS2D, n, n, 1, 1.0, 0, 1, 1 / 5.0, 0.01, PSG_FORCE_KOLMO);
  psg_ctx *ctx;
  if (psg_create(&c, &ctx)) { check(false, "newton create", 1); return; }
  or_grid g;
  or_grid2(&g, n, n, 1.0);
  std::vector<double> kx(g.nkt), ky(g.nkt);
  std::vector<int> L(g.nkt);
  or_tables(&g, kx.data(), ky.data(), nullptr, L.data());
  V z(2 * g.nkt, 0.0), d(2 * g.nkt);
  or_ic_ns2d_cfg(&g, 9, 0.05, (orc *)d.data());
  for (long i = 0; i < g.nkt; i++) {
    if (kx[i] == 0 && ky[i] == 4) z[2 * i] = -5.0 / 8;
    z[2 * i] += d[2 * i]; z[2 * i + 1] += d[2 * i + 1];
  }
  psg_newton_opts o;
  psg_newton_defaults(PSG_NS2D, &o);
  o.nshift = 0; o.fixed_period = 1; o.jv = jv; o.dt0 = 0.01; o.max_gmres = 12; o.max_newton = 8;
  psg_newton_result r;
  double T = 0.2, sx = 0, sz = 0;
  psg_newton_solve(ctx, z.data(), &T, &sx, &sz, &o, &r);
  psg_destroy(ctx);
  check(r.status == PSG_NEWTON_CONVERGED, jv ? "newton 2d laminar (TL Jv)" : "newton 2d laminar (FD Jv)", r.ner);
}

// the tangent linear about a recorded base (psg_tangent_many) against the
// combined base + perturbation stepper (psg_tangent), FP64 and FP32 records,
// over both time policies; and Newton with the tangent-linear Jv with and
// without the record
static void tangent_rec_case(bool with_newton) {
  const int n = 32, nd = 3;
  psg_config c = cfg3(PSG_NS2D, n, n, 1, 1.0, PSG_FP64, 1, 1 / 40.0, 0.01, PSG_FORCE_KOLMO);
  psg_ctx *ctx;
  if (psg_create(&c, &ctx)) { check(false, "tangent rec create", 1); return; }
  or_grid g;
  or_grid2(&g, n, n, 1.0);
  V z(2 * g.nkt), d((size_t)2 * g.nkt * nd);
  or_ic_ns2d_cfg(&g, 3, 1.0, (orc *)z.data());
  for (int i = 0; i < nd; i++) or_ic_ns2d_cfg(&g, 20 + i, 0.3, (orc *)(d.data() + (size_t)2 * g.nkt * i));
  for (int policy = 0; policy < 2; policy++) {
    const double T = policy ? 0.237 : 0.2;
    V z1(2 * g.nkt), zm(2 * g.nkt), dm(d.size()), d1(2 * g.nkt);
    double worst = 0, worst32 = 0, wz = 0;
    int rc = psg_tangent_many(ctx, z.data(), nd, d.data(), T, 0.01, policy, 0, zm.data(), dm.data());
    for (int i = 0; i < nd && !rc; i++) {
      V di(d.begin() + (size_t)2 * g.nkt * i, d.begin() + (size_t)2 * g.nkt * (i + 1));
      rc = psg_tangent(ctx, z.data(), di.data(), T, 0.01, policy, z1.data(), d1.data());
      V mi(dm.begin() + (size_t)2 * g.nkt * i, dm.begin() + (size_t)2 * g.nkt * (i + 1));
      worst = fmax(worst, rel(mi, d1));
      wz = fmax(wz, rel(zm, z1));
    }
    V d32(d.size());
    if (!rc) rc = psg_tangent_many(ctx, z.data(), nd, d.data(), T, 0.01, policy, 1, nullptr, d32.data());
    worst32 = rel(d32, dm);
    char nm[96];
    snprintf(nm, sizeof nm, "tangent about a recorded base, policy %d (fp64)", policy);
    check(rc == 0 && worst < 1e-12 && wz < 1e-13, nm, worst);
    snprintf(nm, sizeof nm, "tangent about a recorded base, policy %d (fp32 record)", policy);
    check(rc == 0 && worst32 < 1e-4 && worst32 > 0, nm, worst32);
  }
  psg_destroy(ctx);
  if (!with_newton) return;
  // Newton: the recorded and the combined tangent linear take the same steps
  c = cfg3(PSG_NS2D, n, n, 1, 1.0, 0, 1, 1 / 5.0, 0.01, PSG_FORCE_KOLMO);
  if (psg_create(&c, &ctx)) { check(false, "newton rec create", 1); return; }
  std::vector<double> kx(g.nkt), ky(g.nkt);
  std::vector<int> L(g.nkt);
  or_tables(&g, kx.data(), ky.data(), nullptr, L.data());
  V z0(2 * g.nkt, 0.0), dd(2 * g.nkt);
  or_ic_ns2d_cfg(&g, 9, 0.05, (orc *)dd.data());
  for (long i = 0; i < g.nkt; i++) {
    if (kx[i] == 0 && ky[i] == 4) z0[2 * i] = -5.0 / 8;
    z0[2 * i] += dd[2 * i]; z0[2 * i + 1] += dd[2 * i + 1];
  }
  psg_newton_result r[2];
  V zs[2] = {z0, z0};
  for (int k = 0; k < 2; k++) {
    psg_newton_opts o;
    psg_newton_defaults(PSG_NS2D, &o);
    o.nshift = 0; o.fixed_period = 1; o.jv = 1; o.dt0 = 0.01; o.max_gmres = 12; o.max_newton = 8;
    o.tl_record = k;
    double T = 0.2, sx = 0, sz = 0;
    psg_newton_solve(ctx, zs[k].data(), &T, &sx, &sz, &o, &r[k]);
  }
  psg_destroy(ctx);
  check(r[0].status == PSG_NEWTON_CONVERGED && r[1].status == PSG_NEWTON_CONVERGED &&
            r[0].newton_iters == r[1].newton_iters && r[0].gmres_iters == r[1].gmres_iters && rel(zs[1], zs[0]) < 1e-12,
        "newton 2d laminar, TL Jv with and without the record", rel(zs[1], zs[0]));
}

// Boussinesq Newton with both shifts and the period free, two Newton steps
// from a developed 16^3 state (status and residual printed)
static void bous_newton_case(int fixed_period) {
  const int n = 16;
  psg_config c = cfg3(PSG_BOUS3D, n, n, n, 1.0, PSG_FP64, 1, 0.02, 0.01, PSG_FORCE_KOLMO);
  psg_ctx *ctx;
  if (psg_create(&c, &ctx)) { check(false, "bous newton create", 1); return; }
  V z = ic3(n, n, n, 1.0, 73, 4);
  {
    or_grid g;
    or_grid3(&g, n, n, n, 1.0);
    orb_ctx *o = orb_create(n, n, n, 1.0, c.v2, c.sc, c.ri, c.theta, 0.01, c.kf, c.ampfor);
    orc *w = (orc *)z.data();
    orb_run(o, w, w + g.nkt, w + 2 * g.nkt, w + 3 * g.nkt, 30);
    orb_destroy(o);
  }
  psg_newton_opts o;
  psg_newton_defaults(PSG_BOUS3D, &o);
  o.fixed_period = fixed_period; o.dt0 = 0.01; o.max_gmres = 30; o.max_newton = 2; o.verbose = 2;
  psg_newton_result r;
  double T = 0.3, sx = 0.05, sz = -0.05;
  const int rc = psg_newton_solve(ctx, z.data(), &T, &sx, &sz, &o, &r);
  if (rc) printf("bous newton error: %s\n", psg_last_error());
  printf("bous newton: status %d iters %d ner %.6e T %.10f sx %.6f sz %.6f\n", r.status, r.newton_iters, r.ner, T, sx,
         sz);
  psg_destroy(ctx);
  check(rc == 0 && std::isfinite(r.ner), fixed_period ? "bous newton two shifts, fixed period" : "bous newton two shifts and period", r.ner);
}

// three guesses of different periods on two lanes against one-guess solves
static void newton2d_batch_case() {
  const int n = 32, ng = 3;
  psg_config c = cfg3(PSG_NS2D, n, n, 1, 1.0, 0, 2, 1 / 5.0, 0.01, PSG_FORCE_KOLMO);
  psg_config c1 = cfg3(PSG_NS2D, n, n, 1, 1.0, 0, 1, 1 / 5.0, 0.01, PSG_FORCE_KOLMO);
  psg_ctx *ctx, *ctx1;
  if (psg_create(&c, &ctx) || psg_create(&c1, &ctx1)) { check(false, "newton batch create", 1); return; }
  or_grid g;
  or_grid2(&g, n, n, 1.0);
  std::vector<double> kx(g.nkt), ky(g.nkt);
  std::vector<int> L(g.nkt);
  or_tables(&g, kx.data(), ky.data(), nullptr, L.data());
  V z((size_t)2 * g.nkt * ng, 0.0), z1;
  for (int q = 0; q < ng; q++) {
    V d(2 * g.nkt);
    or_ic_ns2d_cfg(&g, 9 + q, 0.05, (orc *)d.data());
    for (long i = 0; i < g.nkt; i++) {
      double *zz = z.data() + (size_t)q * 2 * g.nkt;
      if (kx[i] == 0 && ky[i] == 4) zz[2 * i] = -5.0 / 8;
      zz[2 * i] += d[2 * i]; zz[2 * i + 1] += d[2 * i + 1];
    }
  }
  z1 = z;
  psg_newton_opts o;
  psg_newton_defaults(PSG_NS2D, &o);
  o.nshift = 0; o.fixed_period = 1; o.jv = 0; o.dt0 = 0.01; o.max_gmres = 12; o.max_newton = 8;
  psg_newton_result r[ng], r1;
  double T[ng] = {0.2, 0.33, 0.25}, sx[ng] = {0, 0, 0};
  const int rc = psg_newton_batch(ctx, ng, z.data(), T, sx, nullptr, &o, r);
  double worst = 0.0;
  bool same = rc == 0;
  for (int q = 0; q < ng; q++) {
    double T1 = T[q], sx1 = 0, sz1 = 0;
    psg_newton_solve(ctx1, z1.data() + (size_t)q * 2 * g.nkt, &T1, &sx1, &sz1, &o, &r1);
    same = same && r[q].status == PSG_NEWTON_CONVERGED && r1.status == PSG_NEWTON_CONVERGED &&
           r[q].newton_iters == r1.newton_iters && r[q].gmres_iters == r1.gmres_iters;
    for (long i = 0; i < 2 * g.nkt; i++)
      worst = fmax(worst, fabs(z[(size_t)q * 2 * g.nkt + i] - z1[(size_t)q * 2 * g.nkt + i]));
  }
  psg_destroy(ctx);
  psg_destroy(ctx1);
  check(same && worst < 1e-11, "newton 2d batch of 3 guesses on 2 lanes (FD Jv)", worst);
}

// Floquet analysis of the laminar Kolmogorov state (32^2, Re = 40, T = 0.2) with
// the tangent-linear Jv and the Arnoldi size of the 2D reference (ncv = 200):
// runs to completion, leading multiplier real and unstable
static void floquet_case(int ncv, int jv) {
  psg_config c = cfg3(PSG_NS2D, 32, 32, 1, 1.0, PSG_FP64, 1, 1 / 40.0, 0.01, PSG_FORCE_KOLMO);
  psg_ctx *ctx;
  if (psg_create(&c, &ctx)) { check(false, "floquet create", 1); return; }
  or_grid g;
  or_grid2(&g, 32, 32, 1.0);
  std::vector<double> kx(g.nkt), ky(g.nkt), kz(g.nkt);
  std::vector<int> L(g.nkt);
  or_tables(&g, kx.data(), ky.data(), kz.data(), L.data());
  V z(2 * g.nkt, 0.0);
  for (long i = 0; i < g.nkt; i++)
    if (kx[i] == 0.0 && ky[i] == 4.0) z[2 * i] = -2.0 / (16 / 40.0);
  psg_floquet_opts o;
  psg_floquet_defaults(&o);
  o.nev = 10; o.ncv = ncv; o.jv = jv; o.dt0 = 0.01; o.verbose = 1;
  double mr[10], mi[10], er[10];
  psg_floquet_result r;
  const int rc = psg_floquet(ctx, z.data(), 0.2, 0.0, 0.0, &o, mr, mi, er, NULL, &r);
  if (rc) printf("floquet error: %s\n", psg_last_error());
  psg_destroy(ctx);
  printf("floquet: nconv %d matvecs %d cycles %d mu0 %.12g%+.12gi\n", r.nconv, r.matvecs, r.cycles, mr[0], mi[0]);
  check(rc == 0 && mr[0] > 1.0 && mi[0] == 0.0, "floquet 2d laminar Kolmogorov (TL Jv, ncv 200)", er[0]);
}

// Recurrence search through the C-ABI against the oracle on a random history
// (the states are set with psg_set_state; no stepping): residual tables of every
// check, check summaries and guess records.
static void rec_case(const char *name, int flav, int n, int rule, int W, int nsx, int nsy, int nsz, double ri,
                     int store, int entries) {
  const bool two = flav == PSG_NS2D;
  psg_config c = cfg3(flav, n, n, two ? 1 : n, 1.0, PSG_FP64, 1, 0.02, 0.01, PSG_FORCE_KOLMO);
  c.ri = ri;
  psg_ctx *ctx;
  if (psg_create(&c, &ctx)) { check(false, name, 1); return; }
  or_grid g;
  if (two) or_grid2(&g, n, n, 1.0);
  else or_grid3(&g, n, n, n, 1.0);
  std::vector<double> kx(g.nkt), ky(g.nkt), kz(g.nkt);
  std::vector<int> L(g.nkt);
  or_tables(&g, kx.data(), ky.data(), kz.data(), L.data());
  std::vector<int8_t> L8(L.begin(), L.end());
  const int nf = two ? 1 : (flav == PSG_BOUS3D ? 4 : 3);
  const long nk = (long)nf * g.nkt;
  std::vector<V> hist(entries, V(2 * nk));
  uint64_t x = 88172645463325252ull;
  for (auto &h : hist)
    for (long i = 0; i < nk; i++) {
      x ^= x << 13; x ^= x >> 7; x ^= x << 17;
      const double a = (double)(x >> 11) / 9007199254740992.0 - 0.5;
      x ^= x << 13; x ^= x >> 7; x ^= x << 17;
      const double b = (double)(x >> 11) / 9007199254740992.0 - 0.5;
      h[2 * i] = a * L[i % g.nkt];
      h[2 * i + 1] = b * L[i % g.nkt];
    }
  // plant near-recurrences: entry e (e % 4 == 1) is entry e-lag shifted by the
  // x-shift 5 of the grid plus 1% noise, so the check that follows it has a
  // small best residual and the next one a large one
  // (the Boussinesq rule skips periods below 0.5, i.e. lags under 5 entries: lag 6 there)
  const double sx5 = 2.0 * M_PI * 5 / nsx;
  const int lag = flav == PSG_BOUS3D ? 6 : 3;
  for (int e = lag + 2; e < entries; e++) {
    if (e % 4 != 1) continue;
    for (long i = 0; i < nk; i++) {
      const double ph = kx[i % g.nkt] * sx5, cr = cos(ph), ci = sin(ph);
      const double ar = hist[e - lag][2 * i], ai = hist[e - lag][2 * i + 1];
      hist[e][2 * i] = cr * ar - ci * ai + 0.01 * hist[e][2 * i];
      hist[e][2 * i + 1] = cr * ai + ci * ar + 0.01 * hist[e][2 * i + 1];
    }
  }
  psg_set_state(ctx, 0, hist[0].data());
  psg_rec_opts o;
  psg_rec_defaults(ctx, &o);
  o.rule = rule; o.store_size = store; o.norm_width = W; o.nshift_x = nsx; o.nshift_y = nsy; o.nshift_z = nsz;
  o.nout = 10; o.threshold = 0.5; o.ri = ri;
  psg_rec *r;
  if (psg_rec_create(ctx, &o, &r)) { printf("rec: %s\n", psg_last_error()); check(false, name, 1); psg_destroy(ctx); return; }
  or_rec *orr = or_rec_new(rule, nf, g.iktx, g.ikty, two ? 1 : g.iktz, g.kty, two ? 0 : g.ktz, 1.0, L8.data(), store,
                           W, nsx, nsy, nsz, 10, 0.01, 0.0, 0.5, ri, 0.02, hist[0].data());
  const int ns = nsx * nsy * nsz;
  double worst = 0.0;
  bool same = true;
  int guesses = 0;
  for (int e = 1; e < entries; e++) {
    psg_set_state(ctx, 0, hist[e].data());
    psg_rec_guess gg;
    psg_rec_check cc;
    V gs(2 * nk), ogs(2 * nk);
    const int a = psg_rec_update(r, 10L * e, 0.01, &gg, gs.data(), &cc);
    double og[10], oc[9];
    const int b = or_rec_update(orr, 10L * e, 0.01, hist[e].data(), og, ogs.data(), oc);
    if (a != b || cc.compared != (int)oc[1] || cc.store != (int)oc[2] || cc.count != (int)oc[3]) same = false;
    if (a == 1) {
      guesses++;
      if (gg.index != (int)og[0] || gg.store != (int)og[1] || fabs(gg.period - og[4]) > 1e-12 || gs != ogs)
        same = false;
    }
    std::vector<double> res((size_t)store * ns), ores((size_t)store * ns);
    std::vector<int> sl(store), osl(store);
    int nj = 0;
    psg_rec_last(r, res.data(), sl.data(), &nj, store);
    const int onj = or_rec_last(orr, ores.data(), osl.data(), store);
    if (nj != onj) same = false;
    for (int j = 0; j < nj && j < onj; j++) {
      if (sl[j] != osl[j]) same = false;
      for (int s2 = 0; s2 < ns; s2++) {
        // compare res^2 = normFN / norm (an O(1) quantity whose rounding is absolute)
        const double a2 = res[(size_t)j * ns + s2], b2 = ores[(size_t)j * ns + s2];
        worst = fmax(worst, fabs(a2 * a2 - b2 * b2));
      }
    }
  }
  psg_rec_destroy(r);
  or_rec_free(orr);
  psg_destroy(ctx);
  printf("%s: %d guesses, decisions %s\n", name, guesses, same ? "equal" : "DIFFER");
  check(same && worst < 1e-13, name, worst);
}

// Run statistics through the C-ABI: sums and averages against the oracle, the
// physical-space pass (both gravity orientations), spectra against the oracle,
// mean fields finite; a bounds and barrier check of every psg_stats kernel.
static void stats_case(const char *name, int flav, int n, double theta) {
  psg_config c = cfg3(flav, n, n, n, 1.0, PSG_FP64, 1, 0.02, 0.004, PSG_FORCE_KOLMO);
  c.theta = theta;
  c.ri = 0.7;
  psg_ctx *ctx;
  if (psg_create(&c, &ctx)) { check(false, name, 1); return; }
  const int nf = flav == PSG_BOUS3D ? 4 : 3;
  or_grid g;
  or_grid3(&g, n, n, n, 1.0);
  std::vector<double> kx(g.nkt), ky(g.nkt), kz(g.nkt);
  std::vector<int> L(g.nkt);
  or_tables(&g, kx.data(), ky.data(), kz.data(), L.data());
  psg_stats *s;
  if (psg_stats_create(ctx, 0, &s)) { check(false, name, 1); psg_destroy(ctx); return; }
  std::vector<orc> t((size_t)nf * g.nkt, orc{0.0, 0.0});
  double worst = 0.0;
  bool fin = true;
  for (int k = 0; k < 3; k++) {
    V z = ic3(n, n, n, 1.0, 300 + k, nf);
    psg_set_state(ctx, 0, z.data());
    double v[16], w[16];
    psg_stats_update(s, k, v);
    const orc *Z = (const orc *)z.data();
    if (nf == 4)
      or_stats_bous(g.nkt, kx.data(), ky.data(), kz.data(), L.data(), Z, Z + g.nkt, Z + 2 * g.nkt, Z + 3 * g.nkt,
                    t.data(), t.data() + g.nkt, t.data() + 2 * g.nkt, t.data() + 3 * g.nkt, k, c.ampfor, c.kf, 0.7,
                    theta, w);
    else
      or_stats_ns3d(g.nkt, kx.data(), ky.data(), kz.data(), L.data(), Z, Z + g.nkt, Z + 2 * g.nkt, t.data(),
                    t.data() + g.nkt, t.data() + 2 * g.nkt, k, c.ampfor, c.kf, w);
    for (int q = 0; q < (nf == 4 ? (int)PSG_STB_N : (int)PSG_STNS_N); q++)
      worst = fmax(worst, fabs(v[q] - w[q]) / fmax(fabs(w[q]), 1e-300));
    if (nf == 4) {
      double um = 0, na[64], sa[64];
      int np = 0;
      if (psg_stats_physical(s, PSG_STP_RMS | PSG_STP_UMAX | PSG_STP_RIG, k, &um, na, sa, &np) || np != n) fin = false;
      for (int i = 0; i < np; i++) fin = fin && std::isfinite(na[i]) && std::isfinite(sa[i]);
    }
  }
  if (nf == 4) {
    std::vector<double> h(n + 1), vv(n + 1), oh(n + 1), ov(n + 1);
    int size = 0;
    psg_stats_spectra(s, h.data(), vv.data(), n + 1, &size);
    const int os = or_spectra(g.nkt, n, kx.data(), ky.data(), kz.data(), t.data(), t.data() + g.nkt,
                              t.data() + 2 * g.nkt, theta, oh.data(), ov.data());
    if (os != size) fin = false;
    for (int i = 0; i <= n; i++) worst = fmax(worst, fabs(h[i] - oh[i]) / fmax(fabs(oh[i]), 1e-300) * (oh[i] != 0) +
                                                      fabs(vv[i] - ov[i]) / fmax(fabs(ov[i]), 1e-300) * (ov[i] != 0));
    std::vector<double> m((size_t)7 * n * n * n);
    if (psg_stats_mean_fields(s, m.data())) fin = false;
    for (double x : m) fin = fin && std::isfinite(x);
  }
  psg_stats_destroy(s);
  psg_destroy(ctx);
  check(fin && worst < 1e-12, name, worst);
}

int main(int argc, char **argv) {
  omp_set_dynamic(0);
  omp_set_max_active_levels(1);
  const bool quick = argc > 1 && strcmp(argv[1], "quick") == 0;
  if (argc > 1 && strcmp(argv[1], "big2d") == 0) {
    // 2D grids whose x-rows need more than one wavefront per team (N/8 > 64):
    // the split-row x-pass then synchronises its teams with block barriers
    ns2d_case("ns2d 1024^2 fp64", 1024, PSG_FP64, 1, 0, 1);
    printf("%d passed, %d failed\n", g_pass, g_fail);
    return g_fail ? 1 : 0;
  }
  if (argc > 1 && strcmp(argv[1], "floquet") == 0) {
    floquet_case(200, 1);
    printf("%d passed, %d failed\n", g_pass, g_fail);
    return g_fail ? 1 : 0;
  }
  if (argc > 1 && strcmp(argv[1], "bousvar") == 0) {
    for (int v = 0; v < 12; v++) {
      char nm[96];
      snprintf(nm, sizeof nm, "bous 16^3 fp64 xpass variant %d", v);
      bous_case(nm, 16, 16, 16, 1.0, PSG_FP64, 1, 0, v);
      snprintf(nm, sizeof nm, "bous 32x16x16 a=0.5 fp32 xpass variant %d batch 2", v);
      bous_case(nm, 32, 16, 16, 0.5, PSG_FP32BULK, 2, 0, v);
    }
    printf("%d passed, %d failed\n", g_pass, g_fail);
    return g_fail ? 1 : 0;
  }
  if (argc > 1 && strcmp(argv[1], "bousnewton") == 0) {
    bous_newton_case(0);
    printf("%d passed, %d failed\n", g_pass, g_fail);
    return g_fail ? 1 : 0;
  }
  if (argc > 1 && strcmp(argv[1], "newton") == 0) {
    newton2d_case(0);
    newton2d_case(1);
    newton2d_batch_case();
    tangent_rec_case(true);
    printf("%d passed, %d failed\n", g_pass, g_fail);
    return g_fail ? 1 : 0;
  }
  if (argc > 1 && strcmp(argv[1], "stats") == 0) {
    stats_case("stats ns3d 16^3", PSG_NS3D, 16, M_PI / 2);
    stats_case("stats bous 16^3 theta pi/2", PSG_BOUS3D, 16, M_PI / 2);
    stats_case("stats bous 16^3 theta pi/8", PSG_BOUS3D, 16, M_PI / 8);
    printf("%d passed, %d failed\n", g_pass, g_fail);
    return g_fail ? 1 : 0;
  }
  if (argc > 1 && strcmp(argv[1], "rec") == 0) {
    rec_case("rec ns3d 32^3 W 8 shifts 30x2x2 ring 6", PSG_NS3D, 32, PSG_REC_NS, 8, 30, 2, 2, 0.0, 6, 14);
    rec_case("rec bous 32^3 W 8 shifts 64 ring 40", PSG_BOUS3D, 32, PSG_REC_BOUS, 8, 64, 1, 1, 0.7, 40, 12);
    rec_case("rec bous 32^3 W 8 shifts 64 ring 8 (wraps)", PSG_BOUS3D, 32, PSG_REC_BOUS, 8, 64, 1, 1, 0.7, 8, 18);
    rec_case("rec ns2d 128^2 W 42 shifts 30x4 ring 8", PSG_NS2D, 128, PSG_REC_NS2D, 42, 30, 4, 1, 0.0, 8, 12);
    printf("%d passed, %d failed\n", g_pass, g_fail);
    return g_fail ? 1 : 0;
  }
  if (argc > 1 && strcmp(argv[1], "fp32") == 0) {
    // FP32-bulk forms only, at 16^3
    const int vs[] = {0, 1, 3, 8, 10};
    for (int v : vs) {
      char nm[96];
      snprintf(nm, sizeof nm, "ns3d 16^3 fp32 xpass variant %d", v);
      ns3d_case(nm, PSG_NS3D, 16, 16, 16, 1.0, PSG_FP32BULK, v, 0, 2, PSG_FORCE_KOLMO, 1e-4);
    }
    ns3d_case("ns3d 16^3 fp32 zfwd variant 1 graph", PSG_NS3D, 16, 16, 16, 1.0, PSG_FP32BULK, 0, 2, 5,
              PSG_FORCE_KOLMO, 1e-4, 1);
    ns3d_case("euler 16^3 fp32", PSG_EULER3D, 16, 16, 16, 1.0, PSG_FP32BULK, 0, 0, 3, PSG_FORCE_NONE, 1e-4);
    tladj_case(16, PSG_FP32BULK);
    printf("%d passed, %d failed\n", g_pass, g_fail);
    return g_fail ? 1 : 0;
  }
  // 3D Navier-Stokes: forcings, precisions, every x-pass variant, graph path, aspect ratio
  for (int v = 0; v < 12; v++) {
    char nm[96];
    snprintf(nm, sizeof nm, "ns3d 16^3 fp64 xpass variant %d", v);
    ns3d_case(nm, PSG_NS3D, 16, 16, 16, 1.0, PSG_FP64, v, 0, 2, PSG_FORCE_KOLMO, 1e-11);
    if (quick) continue;
    snprintf(nm, sizeof nm, "ns3d 32^3 fp32 xpass variant %d", v);
    ns3d_case(nm, PSG_NS3D, 32, 32, 32, 1.0, PSG_FP32BULK, v, 0, 2, PSG_FORCE_KOLMO, 1e-4);
  }
  ns3d_case("ns3d 16^3 fp64 zfwd variant 1", PSG_NS3D, 16, 16, 16, 1.0, PSG_FP64, 0, 0, 2, PSG_FORCE_KOLMO, 1e-11, 1);
  // passes 2-4 by slabs of z planes: 16 and 8 planes of 32, batches of 2, graph path
  ns3d_case("ns3d 32^3 fp64 slabs of 16", PSG_NS3D, 32, 32, 32, 1.0, PSG_FP64, 1, 0, 2, PSG_FORCE_KOLMO, 1e-11, 0, 4);
  ns3d_case("ns3d 32^3 fp64 slabs of 8 batch 2 graph", PSG_NS3D, 32, 32, 32, 1.0, PSG_FP64, 0, 2, 5, PSG_FORCE_KOLMO,
            1e-11, 1, 5, 2);
  ns3d_case("ns3d 32x16x16 a=0.5 fp32 slabs of 8 batch 3", PSG_NS3D, 32, 16, 16, 0.5, PSG_FP32BULK, 10, 0, 3,
            PSG_FORCE_GMRES, 1e-4, 0, 5, 3);
  ns3d_case("ns3d 32^3 fp64 graph 2 (dns forcing)", PSG_NS3D, 32, 32, 32, 1.0, PSG_FP64, 1, 2, 6, PSG_FORCE_DNS, 1e-11);
  ns3d_case("ns3d 32x16x16 a=0.5 fp64 graph (gmres forcing)", PSG_NS3D, 32, 16, 16, 0.5, PSG_FP64, 0, 2, 5,
            PSG_FORCE_GMRES, 1e-11);
  ns3d_case("euler 16^3 fp64 graph", PSG_EULER3D, 16, 16, 16, 1.0, PSG_FP64, 0, 2, 5, PSG_FORCE_NONE, 1e-11);
  bous_case("bous 16^3 fp64 batch 2 graph", 16, 16, 16, 1.0, PSG_FP64, 2, 2);
  for (int v = 1; v < 12; v++) {
    char nm[96];
    snprintf(nm, sizeof nm, "bous 16^3 fp64 xpass variant %d", v);
    bous_case(nm, 16, 16, 16, 1.0, PSG_FP64, 1, 0, v);
    if (quick) continue;
    snprintf(nm, sizeof nm, "bous 32x16x16 a=0.5 fp32 xpass variant %d batch 2", v);
    bous_case(nm, 32, 16, 16, 0.5, PSG_FP32BULK, 2, 0, v);
  }
  if (!quick) bous_case("bous 32x16x16 a=0.5 fp32", 32, 16, 16, 0.5, PSG_FP32BULK, 1, 0);
  ns2d_case("ns2d 32^2 fp64 graph", 32, PSG_FP64, 1, 2, 5);
  ns2d_case("ns2d 64^2 fp64 batch 3", 64, PSG_FP64, 3, 0, 3);
  if (!quick) ns2d_case("ns2d 64^2 fp32 graph", 64, PSG_FP32BULK, 1, 2, 5);
  tladj_case(16, PSG_FP64);
  if (!quick) tladj_case(16, PSG_FP32BULK);
  newton2d_case(0);
  newton2d_case(1);
  newton2d_batch_case();
  tangent_rec_case(!quick);   // the two emulated Newton solves take most of an hour
  rec_case("rec ns3d 32^3 W 8 shifts 30x2x2 ring 6", PSG_NS3D, 32, PSG_REC_NS, 8, 30, 2, 2, 0.0, 6, 14);
  rec_case("rec bous 32^3 W 8 shifts 64 ring 8 (wraps)", PSG_BOUS3D, 32, PSG_REC_BOUS, 8, 64, 1, 1, 0.7, 8, 18);
  rec_case("rec ns2d 128^2 W 42 shifts 30x4 ring 8", PSG_NS2D, 128, PSG_REC_NS2D, 42, 30, 4, 1, 0.0, 8, 12);
  stats_case("stats ns3d 16^3", PSG_NS3D, 16, M_PI / 2);
  stats_case("stats bous 16^3 theta pi/2", PSG_BOUS3D, 16, M_PI / 2);
  stats_case("stats bous 16^3 theta pi/8", PSG_BOUS3D, 16, M_PI / 8);
  printf("%d passed, %d failed\n", g_pass, g_fail);
  return g_fail ? 1 : 0;
}
