// The C++ host boundary as an escript adapter uses it (include/ripley_hip.hpp
// over the C ABI, host arrays only): the element loop of
// RipleyDomain::addToSystemMatrix inserting host element matrices, the
// constraint of insertConstraint, setToSolution with PASO's solver options.
//
// Problem: Brick(NE0, NE1, NE2) of unit length, -div grad u = 1, u = 0 on
// z = 0, natural conditions elsewhere.  The solution depends on z only and
// the Q1 solution is nodally exact for it: u(z) = z - z^2 / 2.  Checks:
//   * patch test before constraints: (A u_lin)_i = 0 for interior nodes of
//     u_lin = 1 + 2x + 3y + 4z (ypAx, host arrays);
//   * PASO PCG (Jacobi), PCG with Gauss-Seidel and BiCGStab reach the exact
//     nodal values to 1e-7.
// Exit status 0 when every check passes.
#include <array>
#include <cmath>
#include <cstdio>
#include <vector>

#include "ripley_hip.hpp"

namespace {

constexpr int NE0 = 12, NE1 = 10, NE2 = 16;

// Q1 stiffness matrix of one hexahedron (sizes h) for -div grad, by the
// 2x2x2 Gauss rule; local node a has offsets (a & 1, (a >> 1) & 1, a >> 2).
std::array<double, 64> element_stiffness(const double h[3])
{
    const double g[2] = {0.5 - 0.5 / std::sqrt(3.0), 0.5 + 0.5 / std::sqrt(3.0)};
    std::array<double, 64> K{};
    const double w = h[0] * h[1] * h[2] / 8.0;
    for (int q = 0; q < 8; ++q) {
        const double xi[3] = {g[q & 1], g[(q >> 1) & 1], g[q >> 2]};
        double grad[8][3];
        for (int a = 0; a < 8; ++a) {
            double f[3], df[3];
            for (int d = 0; d < 3; ++d) {
                const int bit = (a >> d) & 1;
                f[d] = bit ? xi[d] : 1.0 - xi[d];
                df[d] = (bit ? 1.0 : -1.0) / h[d];
            }
            grad[a][0] = df[0] * f[1] * f[2];
            grad[a][1] = f[0] * df[1] * f[2];
            grad[a][2] = f[0] * f[1] * df[2];
        }
        for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b)
                K[a + 8 * b] += w * (grad[a][0] * grad[b][0] + grad[a][1] * grad[b][1] + grad[a][2] * grad[b][2]);
    }
    return K;
}

int failures = 0;

void expect(bool ok, const char* what, double value)
{
    std::printf("%s %s (%.3e)\n", ok ? "ok  " : "FAIL", what, value);
    if (!ok) ++failures;
}

}  // namespace

int main()
{
    const int nn[3] = {NE0 + 1, NE1 + 1, NE2 + 1};
    const int64_t N = (int64_t)nn[0] * nn[1] * nn[2];
    const double h[3] = {1.0 / NE0, 1.0 / NE1, 1.0 / NE2};
    const std::array<double, 64> K = element_stiffness(h);

    ripley_hip::Context ctx(0);
    ripley_hip::SystemMatrix A(ctx, 3, {NE0, NE1, NE2}, {1.0, 1.0, 1.0}, 1, false);

    // element loop (addToSystemMatrix), all elements in one batch
    const int64_t E = (int64_t)NE0 * NE1 * NE2;
    std::vector<int32_t> nodes(E * 8);
    std::vector<double> em(E * 64);
    std::vector<double> rhs(N, 0.0);
    for (int k2 = 0; k2 < NE2; ++k2)
        for (int k1 = 0; k1 < NE1; ++k1)
            for (int k0 = 0; k0 < NE0; ++k0) {
                const int64_t e = k0 + (int64_t)NE0 * (k1 + (int64_t)NE1 * k2);
                const int64_t first = k0 + (int64_t)nn[0] * (k1 + (int64_t)nn[1] * k2);
                for (int a = 0; a < 8; ++a) {
                    const int64_t node = first + (a & 1) + ((a >> 1) & 1) * nn[0] + (a >> 2) * nn[0] * nn[1];
                    nodes[e * 8 + a] = (int32_t)node;
                    rhs[node] += h[0] * h[1] * h[2] / 8.0;   // Y = 1: each shape function integrates to vol/8
                }
                for (int i = 0; i < 64; ++i) em[e * 64 + i] = K[i];
            }
    A.add(E, 8, nodes.data(), em.data());

    auto coord = [&](int64_t i, int d) {
        const int64_t c[3] = {i % nn[0], (i / nn[0]) % nn[1], i / ((int64_t)nn[0] * nn[1])};
        return c[d] * h[d];
    };

    // patch test: A u_lin vanishes on interior rows
    std::vector<double> ul(N), y(N, 0.0);
    for (int64_t i = 0; i < N; ++i) ul[i] = 1.0 + 2.0 * coord(i, 0) + 3.0 * coord(i, 1) + 4.0 * coord(i, 2);
    A.ypAx(y.data(), ul.data());
    double worst = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        const int64_t c0 = i % nn[0], c1 = (i / nn[0]) % nn[1], c2 = i / ((int64_t)nn[0] * nn[1]);
        if (c0 > 0 && c0 < NE0 && c1 > 0 && c1 < NE1 && c2 > 0 && c2 < NE2) worst = std::fmax(worst, std::fabs(y[i]));
    }
    expect(worst <= 1e-12, "patch test: interior rows of A u_lin vanish", worst);

    // u = 0 on z = 0 (insertConstraint with r = 0)
    std::vector<double> q(N, 0.0);
    for (int64_t i = 0; i < N; ++i)
        if (i < (int64_t)nn[0] * nn[1]) {
            q[i] = 1.0;
            rhs[i] = 0.0;
        }
    A.nullifyRowsAndCols(q.data(), q.data(), 1.0);

    struct Case {
        const char* name;
        int method, preconditioner;
    };
    const Case cases[] = {{"PASO PCG, Jacobi", RG_METHOD_PCG, RG_PRECONDITIONER_JACOBI},
                          {"PASO PCG, Gauss-Seidel", RG_METHOD_PCG, RG_PRECONDITIONER_GAUSS_SEIDEL},
                          {"BiCGStab, Jacobi", RG_METHOD_BICGSTAB, RG_PRECONDITIONER_JACOBI}};
    for (const Case& c : cases) {
        ripley_hip::SolverOptions o;
        o.tolerance = 1e-10;
        o.method = c.method;
        o.preconditioner = c.preconditioner;
        std::vector<double> u(N, 0.0);
        const ripley_hip::Diagnostics d = A.setToSolution(u.data(), rhs.data(), o);
        double err = 0.0;
        for (int64_t i = 0; i < N; ++i) {
            const double z = coord(i, 2);
            err = std::fmax(err, std::fabs(u[i] - (z - 0.5 * z * z)));
        }
        std::printf("     %s: %d iterations\n", c.name, d.num_iter);
        expect(d.converged != 0 && err <= 1e-7, c.name, err);
    }
    return failures == 0 ? 0 : 1;
}
