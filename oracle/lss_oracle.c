/*
 * lss_oracle.c -- CPU restatement of the reference's host Heston/SABR ADI path.
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg; never by the product library.  See lss_oracle.h for the reference files it follows.
 */
#include "lss_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* grids of at least this many points run the lines of each sweep on OpenMP threads */
static int64_t lss_or_parallel_min = 262144;
void lss_or_set_parallel_min(int64_t points) { lss_or_parallel_min = points; }

/* ---- double instantiation ---- */
#define FP double
#define SFX _f64
#define SINH sinh
#define COSH cosh
#define ASINH asinh
#include "lss_oracle_impl.inc"
#include "lss_oracle_heat1d.inc"
#undef FP
#undef SFX
#undef SINH
#undef COSH
#undef ASINH

/* ---- float instantiation (std::sinh(float) etc. resolve to the float overloads) ---- */
#define FP float
#define SFX _f32
#define SINH sinhf
#define COSH coshf
#define ASINH asinhf
#include "lss_oracle_impl.inc"
#include "lss_oracle_heat1d.inc"
#undef FP
#undef SFX
#undef SINH
#undef COSH
#undef ASINH

/* ------------------------------------------------------------------------------------------------
 * Heston data as written by unit_tests/common/lss_print_t.hpp:2156-2186 (and the builder test
 * unit_tests/builders/lss_2d_general_heston_equation_builder_t.hpp:68-74).  The captured constants
 * are double, so every expression is evaluated in double, left to right.
 * ---------------------------------------------------------------------------------------------- */
static double hes_a(void *c, double t, double s, double v) { (void)c; (void)t; return (0.5 * v * s * s); }
static double hes_b(void *c, double t, double s, double v)
{
    const lss_or_heston *p = (const lss_or_heston *)c;
    (void)t; (void)s;
    return (0.5 * p->sig_sig * p->sig_sig * v);
}
static double hes_c(void *c, double t, double s, double v)
{
    const lss_or_heston *p = (const lss_or_heston *)c;
    (void)t;
    return (p->rho * p->sig_sig * v * s);
}
static double hes_d(void *c, double t, double s, double v)
{
    const lss_or_heston *p = (const lss_or_heston *)c;
    (void)t; (void)v;
    return (p->rate * s);
}
static double hes_e(void *c, double t, double s, double v)
{
    const lss_or_heston *p = (const lss_or_heston *)c;
    (void)t; (void)s;
    return (p->sig_kappa * (p->sig_theta - v));
}
static double hes_f(void *c, double t, double s, double v)
{
    const lss_or_heston *p = (const lss_or_heston *)c;
    (void)t; (void)s; (void)v;
    return (-p->rate);
}
/* terminal condition std::max<T>(0.0, s - strike): s - strike in double, rounded to T by the
 * caller's cast; max(0, v) keeps 0 when v <= 0.  Rounding happens in the oracle (FP cast), and
 * max commutes with that rounding. */
static double hes_init(void *c, double s, double v)
{
    const lss_or_heston *p = (const lss_or_heston *)c;
    (void)v;
    const double d = s - p->strike;
    return d > 0.0 ? d : 0.0;
}
static double bnd_low(void *c, double t, double y)
{
    (void)t; (void)y;
    return ((const lss_or_heston *)c)->bnd_low;
}
static double bnd_high(void *c, double t, double y)
{
    (void)t; (void)y;
    return ((const lss_or_heston *)c)->bnd_high;
}
static double bnd_identity(void *c, double t, double s) { (void)c; (void)t; return s; }

void lss_or_heston_pde(const lss_or_heston *p, lss_or_pde *out)
{
    out->a = hes_a;
    out->b = hes_b;
    out->c = hes_c;
    out->d = hes_d;
    out->e = hes_e;
    out->f = hes_f;
    out->init = hes_init;
    out->h_lower = bnd_low;
    out->h_upper = bnd_high;
    out->v_upper = bnd_identity;
    out->src = NULL;
    out->ctx = (void *)p;
}

/* ------------------------------------------------------------------------------------------------
 * SABR data as written by unit_tests/common/lss_print_t.hpp:2039-2050:
 * D = exp(-rate*(T - t)); a = 0.5 a^2 s^(2b) D^(2(1-b)); b = 0.5 sig^2 a^2;
 * c = rho sig a^2 s^b D^(1-b); d = rate s; e = 0; f = -rate.
 * ---------------------------------------------------------------------------------------------- */
static double sabr_D(const lss_or_sabr *p, double t) { return exp(-p->rate * (p->maturity - t)); }
static double sabr_a(void *c, double t, double s, double al)
{
    const lss_or_sabr *p = (const lss_or_sabr *)c;
    return (0.5 * al * al * pow(s, 2.0 * p->beta) * pow(sabr_D(p, t), 2.0 * (1.0 - p->beta)));
}
static double sabr_b(void *c, double t, double s, double al)
{
    const lss_or_sabr *p = (const lss_or_sabr *)c;
    (void)t; (void)s;
    return (0.5 * p->sig_sig * p->sig_sig * al * al);
}
static double sabr_c(void *c, double t, double s, double al)
{
    const lss_or_sabr *p = (const lss_or_sabr *)c;
    return (p->rho * p->sig_sig * al * al * pow(s, p->beta) * pow(sabr_D(p, t), (1.0 - p->beta)));
}
static double sabr_d(void *c, double t, double s, double al)
{
    (void)t; (void)al;
    return (((const lss_or_sabr *)c)->rate * s);
}
static double sabr_e(void *c, double t, double s, double al) { (void)c; (void)t; (void)s; (void)al; return 0.0; }
static double sabr_f(void *c, double t, double s, double al)
{
    (void)t; (void)s; (void)al;
    return (-((const lss_or_sabr *)c)->rate);
}
static double sabr_init(void *c, double s, double al)
{
    (void)al;
    const double d = s - ((const lss_or_sabr *)c)->strike;
    return d > 0.0 ? d : 0.0;
}
static double sabr_low(void *c, double t, double y) { (void)t; (void)y; return ((const lss_or_sabr *)c)->bnd_low; }
static double sabr_high(void *c, double t, double y) { (void)t; (void)y; return ((const lss_or_sabr *)c)->bnd_high; }

void lss_or_sabr_pde(const lss_or_sabr *p, lss_or_pde *out)
{
    out->a = sabr_a;
    out->b = sabr_b;
    out->c = sabr_c;
    out->d = sabr_d;
    out->e = sabr_e;
    out->f = sabr_f;
    out->init = sabr_init;
    out->h_lower = sabr_low;
    out->h_upper = sabr_high;
    out->v_upper = bnd_identity;
    out->src = NULL;
    out->ctx = (void *)p;
}

/* ------------------------------------------------------------------------------------------------
 * Black-Scholes data of unit_tests/pde_solvers/one_dimensional/lss_black_scholes_equation_t.hpp:66-80
 * (and unit_tests/common/lss_print_t.hpp:151-166): the captured constants are double, so each
 * expression is evaluated in double.
 * ---------------------------------------------------------------------------------------------- */
static double bs_a(void *c, double t, double x)
{
    const lss_or_bs *p = (const lss_or_bs *)c;
    (void)t;
    return 0.5 * p->sig * p->sig * x * x;
}
static double bs_b(void *c, double t, double x)
{
    const lss_or_bs *p = (const lss_or_bs *)c;
    (void)t;
    return p->rate * x;
}
static double bs_c(void *c, double t, double x)
{
    const lss_or_bs *p = (const lss_or_bs *)c;
    (void)t; (void)x;
    return -p->rate;
}
static double bs_init(void *c, double x)
{
    const lss_or_bs *p = (const lss_or_bs *)c;
    const double v = x - p->strike;
    return v > 0.0 ? v : 0.0;
}
static double bs_lo(void *c, double t) { (void)c; (void)t; return 0.0; }
static double bs_hi(void *c, double t)
{
    const lss_or_bs *p = (const lss_or_bs *)c;
    return (p->x_hi - p->strike * exp(-p->rate * (p->maturity - t)));
}

void lss_or_bs_heat1d(const lss_or_bs *p, lss_or_heat1d_pde *out)
{
    memset(out, 0, sizeof(*out));
    out->a = bs_a;
    out->b = bs_b;
    out->c = bs_c;
    out->src = NULL;
    out->init = bs_init;
    out->lo_type = LSS_OR_BC_DIRICHLET;
    out->hi_type = LSS_OR_BC_DIRICHLET;
    out->lo_val = bs_lo;
    out->hi_val = bs_hi;
    out->ctx = (void *)p;
}
