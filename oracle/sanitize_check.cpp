// Host sanitizer run of the oracle (test infrastructure): every entry point
// of oracle.h on a small mixed system (two particle types, mixed sigma,
// exclusion groups, bonds, CST elements of each material model), built
// with -fsanitize=address,undefined by `make -C oracle sanitize`.  Exits 0
// when every call returned normally; the sanitizers abort on the first
// invalid access, overflow or undefined operation.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "oracle.h"

int main() {
    const int n = 600, ntri = 40, nb = 30;
    const double L = 30.0;
    std::mt19937 rng(7);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::vector<double> x(n), y(n), sig(n), rx(n), ry(n);
    std::vector<int32_t> typ(n), excl(n, -1);
    const int m = 25;  // 25 x 24 lattice, spacing 1.2
    for (int i = 0; i < n; ++i) {
        x[i] = (i % m + 0.5) * (L / m) + 0.1 * (u(rng) - 0.5);
        y[i] = (i / m + 0.5) * (L / 24) + 0.1 * (u(rng) - 0.5);
        sig[i] = u(rng) < 0.5 ? 0.9 : 1.1;
        typ[i] = u(rng) < 0.5 ? 0 : 1;
        rx[i] = x[i];
        ry[i] = y[i];
    }
    for (int i = 0; i < 60; ++i) excl[i] = i / 20;
    std::vector<int32_t> tri, bi, bj;
    for (int e = 0; e < ntri; ++e) {  // triangles of neighbouring lattice sites
        const int a = (e / 20) * m + e % 20, b = a + 1, c = a + m;  // inside rows 0-1
        tri.insert(tri.end(), {a, b, c});
    }
    std::vector<double> tE(ntri, 5.0), tnu(ntri, 0.3), bk, br0;
    for (int k = 0; k < nb; ++k) {
        bi.push_back(100 + k);
        bj.push_back(101 + k);
        bk.push_back(50.0);
        br0.push_back(1.2);
    }
    const double eps[4] = {1.0, 0.6, 0.6, 1.3}, rc[4] = {2.5, 2.0, 2.0, 2.2},
                 ush[4] = {0.0, 0.02, 0.02, 0.0};
    or_system s{};
    s.n = n;
    s.Lx = L;
    s.Ly = L;
    s.x = x.data();
    s.y = y.data();
    s.sigma = sig.data();
    s.type = typ.data();
    s.excl = excl.data();
    s.ntypes = 2;
    s.eps = eps;
    s.rc = rc;
    s.ushift = ush;
    s.lj_on = 1;
    s.nbond = nb;
    s.bi = bi.data();
    s.bj = bj.data();
    s.bk = bk.data();
    s.br0 = br0.data();
    s.ntri = ntri;
    s.tri = tri.data();
    s.rx = rx.data();
    s.ry = ry.data();
    s.tE = tE.data();
    s.tnu = tnu.data();

    std::vector<float> xf(n), yf(n);
    for (int i = 0; i < n; ++i) {
        xf[i] = (float)x[i];
        yf[i] = (float)y[i];
    }
    or_wrap_f32(n, xf.data(), yf.data(), (float)L, (float)L);
    const int ncx = or_cell_count(L, 2.75);
    std::vector<int32_t> key(n), start(ncx * ncx + 1), perm(n);
    or_cell_keys_f32(n, xf.data(), yf.data(), (float)L, (float)L, ncx, ncx, key.data());
    or_counting_sort(n, key.data(), ncx * ncx, start.data(), perm.data());
    std::vector<int32_t> cnt(n);
    std::vector<int64_t> s1(n), s2(n);
    for (int mode = 0; mode < 2; ++mode)
        if (or_pair_sets_f32(&s, mode, cnt.data(), s1.data(), s2.data()) != 0) return 1;
    std::vector<double> fx(n), fy(n), en(OR_NEN), fs(2 * n);
    for (int model = 0; model < 3; ++model) {
        s.model = model;
        for (int p32 = 0; p32 < 2; ++p32)
            for (int all = 0; all < 2; ++all)
                if (or_forces(&s, p32, all, fx.data(), fy.data(), en.data(), fs.data()) != 0)
                    return 2;
    }
    s.model = 0;
    std::vector<double> vx(n), vy(n), mass(n, 1.0), inv_m(n, 1.0), hist((OR_NEN + 2) * 20);
    for (int i = 0; i < n; ++i) {
        vx[i] = 0.3 * (u(rng) - 0.5);
        vy[i] = 0.3 * (u(rng) - 0.5);
    }
    std::vector<double> xr = x, yr = y;
    if (or_run(&s, xr.data(), yr.data(), vx.data(), vy.data(), mass.data(), inv_m.data(), 20,
               0.002, 1, hist.data()) != 0)
        return 3;
    std::vector<float> vxf(n), vyf(n), mf(n, 1.0f);
    for (int i = 0; i < n; ++i) {
        vxf[i] = (float)vx[i];
        vyf[i] = (float)vy[i];
    }
    s.x = x.data();
    s.y = y.data();
    for (int threads = 1; threads <= 2; ++threads) {
        or_baseline *b = or_bl_create(&s, vxf.data(), vyf.data(), mf.data(), 0.002f, 0.4f, threads);
        if (!b) return 4;
        or_bl_run(b, 15);
        std::vector<float> bx(n), by(n), bvx(n), bvy(n);
        double props[5];
        or_bl_get(b, bx.data(), by.data(), bvx.data(), bvy.data(), props);
        or_bl_destroy(b);
        for (int i = 0; i < n; ++i)
            if (!std::isfinite(bx[i]) || !std::isfinite(by[i])) return 5;
    }
    std::printf("oracle sanitizer run ok\n");
    return 0;
}
