/*
 * ORACLE — test infrastructure only. Nothing in the product path may link,
 * load or call this file; it exists so that tests/, the smoke check and the
 * CPU-baseline leg of bench.py can check the HIP path against an independent
 * CPU restatement of the reference algorithm.
 *
 * Plain-C restatement of the Love (1929) cell integral and the
 * translation-invariant influence table of the reference:
 *   asinh_diff      src/boussinesq_kernel.cpp:60-69
 *   love_uz         src/boussinesq_kernel.cpp:44,71-95 (LOVE_FAR_SWITCH 1500)
 *   influence table src/boussinesq_kernel.cpp:97-106
 *                   (T[dy*Ns+dx] = love_uz(dx h, dy h, h/2, h/2) / (pi E*))
 *
 * Pinned by: the reference's 80-digit fixture tests/data/love_reference.txt
 * (472 cases, rel <= 1e-12, copied to tests/golden/), the closed-form self
 * term 4h ln(1+sqrt 2), exact parity, and homogeneity (tests/test_oracle_love.py).
 */
#include <math.h>
#include <stdint.h>

#define LOVE_FAR_SWITCH 1500.0

/* asinh(u) - asinh(v) with the exactly known difference du = u - v.
 * sinh(A-B) identity; conjugate form when u, v share a sign. */
double oracle_asinh_diff(double u, double v, double du) {
    const double huge_arg = 1e150;
    if (fabs(u) > huge_arg || fabs(v) > huge_arg) return asinh(u) - asinh(v);
    const double su = sqrt(1.0 + u * u), sv = sqrt(1.0 + v * v);
    const double den = u * sv + v * su;
    const double w = (u * v > 0.0 && den != 0.0) ? du * (u + v) / den
                                                 : u * sv - v * su;
    return asinh(w);
}

static double term(double t, double s1, double s2, double ds) {
    if (t == 0.0) return 0.0;
    const double c = fabs(t);
    return t * oracle_asinh_diff(s1 / c, s2 / c, ds / c);
}

/* Integral of 1/|r - r'| over the rectangle [-a,a]x[-b,b] at (x, y). */
double oracle_love_uz(double x, double y, double a, double b) {
    x = fabs(x);
    y = fabs(y);
    const double m = (a > b) ? a : b;
    const double r2 = x * x + y * y;
    if (r2 > (LOVE_FAR_SWITCH * m) * (LOVE_FAR_SWITCH * m)) {
        const double r = sqrt(r2);
        const double A = 4.0 * a * b;
        const double quad =
            (a * a * (3.0 * x * x - r2) + b * b * (3.0 * y * y - r2)) /
            (6.0 * r2 * r2);
        return A * (1.0 + quad) / r;
    }
    const double xp = x + a, xm = x - a, yp = y + b, ym = y - b;
    const double da = 2.0 * a, db = 2.0 * b;
    return term(xp, yp, ym, db) + term(yp, xp, xm, da) +
           term(xm, ym, yp, -db) + term(ym, xm, xp, -da);
}

/* Vectorised evaluation for the tests: out[i] = love_uz(x[i], y[i], a[i], b[i]). */
void oracle_love_many(int64_t n, const double* x, const double* y,
                      const double* a, const double* b, double* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = oracle_love_uz(x[i], y[i], a[i], b[i]);
}

/* Ns x Ns table, row-major in dy: T[dy*Ns+dx] = love_uz(dx h, dy h, h/2, h/2)/(pi E*). */
void oracle_love_table(int Ns, double L, double E_star, double* table) {
    const double h = L / Ns;
    const double a = 0.5 * h;
    const double scale = 1.0 / (M_PI * E_star);
    for (int dy = 0; dy < Ns; ++dy)
        for (int dx = 0; dx < Ns; ++dx)
            table[(int64_t)dy * Ns + dx] = scale * oracle_love_uz(dx * h, dy * h, a, a);
}

/* ── Cerruti (tangential) cell integrals: src/cerruti_kernel.cpp:1-147 ──────
 * Unit tangential traction q_x on [-a,a]x[-b,b], surface displacement
 * brackets at (x, y):
 *   uxx = (1-nu) love_uz + nu ylog,  ylog = integral of (x-xi)^2/rho^3
 *   uxy = integral of (x-xi)(y-eta)/rho^3 (odd in x and in y)
 * with the same far-field switch (1500 half-widths) and stable asinh pairs.
 * Pinned by Gauss-Legendre quadrature (tests/test_cerruti.cpp:53-78), the
 * nu = 0 reduction to love_uz, exact parity and the self term
 * (2-nu) 4a ln(1+sqrt 2) (tests/test_oracle.py). */
#define CERRUTI_FAR_SWITCH 1500.0

static double cerruti_ylog(double xa, double ya, double a, double b) {
    const double mmax = (a > b) ? a : b;
    const double r2 = xa * xa + ya * ya;
    if (r2 > (CERRUTI_FAR_SWITCH * mmax) * (CERRUTI_FAR_SWITCH * mmax)) {
        const double r = sqrt(r2);
        const double r3 = r2 * r, r5 = r3 * r2, r7 = r5 * r2;
        const double A = 4.0 * a * b;
        const double x2 = xa * xa, y2 = ya * ya;
        return A * (x2 / r3 + (a * a / 6.0) * (2.0 / r3 - 15.0 * x2 / r5 + 15.0 * x2 * x2 / r7) +
                    (b * b / 6.0) * (-3.0 * x2 / r5 + 15.0 * x2 * y2 / r7));
    }
    const double k = xa + a, l = xa - a, m = ya + b, n = ya - b;
    const double da = 2.0 * a;
    return term(m, k, l, da) + term(n, l, k, -da);
}

double oracle_cerruti_uxx(double x, double y, double a, double b, double nu) {
    const double xa = fabs(x), ya = fabs(y);
    return (1.0 - nu) * oracle_love_uz(xa, ya, a, b) + nu * cerruti_ylog(xa, ya, a, b);
}

double oracle_cerruti_uxy(double x, double y, double a, double b) {
    const double xa = fabs(x), ya = fabs(y);
    const double k = xa + a, l = xa - a, m = ya + b, n = ya - b;
    const double Rkm = hypot(k, m), Rkn = hypot(k, n), Rlm = hypot(l, m), Rln = hypot(l, n);
    const double num = 16.0 * a * b * xa * ya * (1.0 / (Rkn + Rln) + 1.0 / (Rkm + Rlm));
    const double den = (Rlm + Rln) * (Rkn + Rkm);
    const double v = den > 0.0 ? num / den : 0.0;
    return ((x >= 0.0) ? 1.0 : -1.0) * ((y >= 0.0) ? 1.0 : -1.0) * v;
}

void oracle_cerruti_many(int64_t n, const double* x, const double* y, const double* a,
                         const double* b, double nu, double* uxx, double* uxy) {
    for (int64_t i = 0; i < n; ++i) {
        uxx[i] = oracle_cerruti_uxx(x[i], y[i], a[i], b[i], nu);
        uxy[i] = oracle_cerruti_uxy(x[i], y[i], a[i], b[i]);
    }
}

/* CerrutiKernel tables (src/cerruti_kernel.cpp:117-127): xx and xy for
 * dx, dy in [0, Ns), prefactor 1/(pi E* (1-nu)), nu more on xy. */
void oracle_cerruti_tables(int Ns, double L, double E_star, double nu, double* xx, double* xy) {
    const double h = L / Ns, a = 0.5 * h;
    const double pref = 1.0 / (M_PI * E_star * (1.0 - nu)), prefnu = nu * pref;
    for (int dy = 0; dy < Ns; ++dy)
        for (int dx = 0; dx < Ns; ++dx) {
            const int64_t at = (int64_t)dy * Ns + dx;
            xx[at] = pref * oracle_cerruti_uxx(dx * h, dy * h, a, a, nu);
            xy[at] = prefnu * oracle_cerruti_uxy(dx * h, dy * h, a, a);
        }
}
