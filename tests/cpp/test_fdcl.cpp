// test_fdcl.cpp -- checks of the fdcl:: class layer, modelled on the reference's own test
// (test/test.cpp: check_transform round trips at small l_max, 1e-10) and its check_* members.
//
//   test_fdcl cpu         host-only checks (weights, Wigner-d, real harmonics, Legendre, layouts)
//   test_fdcl gpu         transforms through the GPU: round trips, grid vs pointwise inverse
//   test_fdcl dump BETA L print d^l_mn(BETA) for l <= L, packed order, one value per line
//   test_fdcl dumpnp X L  print nP^l_m(X) for l <= L, packed order
//
// Exit status 0 iff every check passes; one line per check on stdout.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "fdcl_FFTS2.hpp"
#include "fdcl_FFTSO3.hpp"

static int failures = 0;

static void expect(const char *name, double err, double tol)
{
    const bool ok = std::isfinite(err) && err <= tol;
    std::printf("%s %s err=%.3e tol=%.1e\n", ok ? "PASS" : "FAIL", name, err, tol);
    if (!ok) failures++;
}

static int cpu_checks()
{
    std::srand(7);
    // packed layout of the containers: F(l,m,n) at off(l) + (m+l) + (n+l)(2l+1); F[l](m+l, n+l)
    {
        fdcl::FFTSO3_matrix_complex F(4);
        double err = 0;
        long e = 0;
        for (int l = 0; l <= 4; l++)
            for (int n = -l; n <= l; n++)
                for (int m = -l; m <= l; m++, e++) {
                    F(l, m, n) = fdcl::complex<double>(l, 10 * m + n);
                    err += std::abs(F.data()[e] - fdcl::complex<double>(l, 10 * m + n));
                    err += std::abs(F[l](m + l, n + l) - F(l, m, n));
                }
        err += std::fabs((double)e - (double)F.size());
        expect("so3_matrix_layout", err, 0);
        fdcl::FFTS2_matrix_real G(5);
        err = 0;
        e = 0;
        for (int l = 0; l <= 5; l++)
            for (int m = -l; m <= l; m++, e++) {
                G(l, m) = 100 * l + m;
                err += std::fabs(G.data()[e] - (100 * l + m)) + std::fabs(G[l](m + l) - G(l, m));
            }
        expect("s2_matrix_layout", err, 0);
    }
    for (int L : {2, 7, 16}) {
        fdcl::FFTSO3_complex fc(L);
        expect(("so3_check_weight_l" + std::to_string(L)).c_str(), fc.check_weight(), 1e-12);
        expect(("so3_check_wigner_d_l" + std::to_string(L)).c_str(), fc.check_wigner_d(), 1e-12);
        fdcl::FFTSO3_real fr(L);
        expect(("so3_check_real_harmonics_l" + std::to_string(L)).c_str(), fr.check_real_harmonics(), 1e-12);
        fdcl::FFTS2_complex sc(L);
        expect(("s2_check_weight_l" + std::to_string(L)).c_str(), sc.check_weight(), 1e-12);
    }
    // the reference's remaining self checks (test/test.cpp:53,57,66,70 hold them below 1e-10)
    for (int L : {5, 20}) {
        fdcl::FFTSO3_complex fc(L);
        expect(("so3_check_deriv_wigner_D_l" + std::to_string(L)).c_str(), fc.check_deriv_wigner_D(), 1e-10);
        fdcl::FFTSO3_real fr(L);
        expect(("so3_check_deriv_real_harmonics_l" + std::to_string(L)).c_str(), fr.check_deriv_real_harmonics(), 1e-10);
    }
    {
        fdcl::FFTSO3_complex fc(5);
        expect("so3_check_Clebsch_Gordon", fc.check_Clebsch_Gordon(), 1e-10);
        fdcl::FFTSO3_real fr(5);
        expect("so3_check_Clebsch_Gordon_real", fr.check_Clebsch_Gordon(), 1e-10);
        // real derivative matrices are real: conj(T) u T^T has no imaginary part; spot values of u_3
        const auto u = fc.deriv_wigner_D();
        expect("so3_deriv_u3_value", std::abs(u[3](4, 3, 3) - fdcl::complex<double>(0, -3)), 0);
        // CG orthogonality: the (l1,l2) block is an orthogonal matrix
        fdcl::Clebsch_Gordon_complex cg(3, 4);
        cg.compute(3, 4);
        const int n = 7 * 9;
        double eo = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                double t = 0;
                for (int q = 0; q < n; q++) t += cg.C(q, i) * cg.C(q, j);
                eo = std::max(eo, std::fabs(t - (i == j ? 1.0 : 0.0)));
            }
        expect("cg_block_orthogonal", eo, 1e-12);
    }
    // d at beta and -beta: d^l_mn(-beta) = (-1)^(m-n) d^l_mn(beta); wigner_D phases
    {
        fdcl::FFTSO3_complex fc(9);
        const double b = 0.7;
        const auto d1 = fc.wigner_d(b), d2 = fc.wigner_d(-b);
        double err = 0;
        for (int l = 0; l <= 9; l++)
            for (int m = -l; m <= l; m++)
                for (int n = -l; n <= l; n++)
                    err = std::max(err, std::fabs(d2(l, m, n) - (((m - n) & 1) ? -1.0 : 1.0) * d1(l, m, n)));
        expect("wigner_d_negative_beta", err, 1e-13);
        const auto D = fc.wigner_D(0.3, b, -1.1);
        err = 0;
        for (int l = 0; l <= 9; l++)
            for (int m = -l; m <= l; m++)
                for (int n = -l; n <= l; n++)
                    err = std::max(err, std::abs(D(l, m, n) - std::exp(-fdcl::I * (0.3 * m - 1.1 * n)) * d1(l, m, n)));
        expect("wigner_D_phases", err, 1e-14);
        // rotation-matrix overloads agree with the Euler-angle ones
        const fdcl::Matrix3d R = fdcl::Euler3232R(0.3, b, 1.9);
        const auto DR = fc.wigner_D(R);
        const auto DE = fc.wigner_D(0.3, b, 1.9);
        expect("wigner_D_rotation_overload", (DR - DE).norm(), 1e-12);
    }
    // copies never share a device plan (the reference's objects are independent values)
    {
        fdcl::FFTSO3_real a(6);
        fdcl::FFTSO3_real b(a);
        fdcl::FFTSO3_real c(3);
        c = a;
        std::vector<fdcl::FFTSO3_complex> v(3, fdcl::FFTSO3_complex(5));
        fdcl::FFTS2_real s(4);
        fdcl::FFTS2_real t(s);
        const bool distinct = a.device_plan_id() != b.device_plan_id() && a.device_plan_id() != c.device_plan_id() &&
                              b.device_plan_id() != c.device_plan_id() && v[0].device_plan_id() != v[1].device_plan_id() &&
                              v[1].device_plan_id() != v[2].device_plan_id() && s.device_plan_id() != t.device_plan_id();
        expect("copies_have_own_device_plans", distinct ? 0.0 : 1.0, 0);
        expect("copy_keeps_l_max", (double)(b.l_max != 6) + (double)(c.l_max != 6) + (double)(t.l_max != 4), 0);
    }
    // pointwise real inverse equals the real part of the complex inverse of T^H F_r T
    {
        const int L = 4;
        fdcl::FFTSO3_real fr(L);
        fdcl::FFTSO3_matrix_real Fr(L);
        Fr.setRandom();
        const auto T = fr.matrix2rsph(L);
        fdcl::FFTSO3_matrix_complex Fc(L);
        for (int l = 0; l <= L; l++)
            for (int m = -l; m <= l; m++)
                for (int n = -l; n <= l; n++) {
                    fdcl::complex<double> s = 0;
                    for (int a = -l; a <= l; a++)
                        for (int c = -l; c <= l; c++) s += std::conj(T(l, a, m)) * Fr(l, a, c) * T(l, c, n);
                    Fc(l, m, n) = s;
                }
        fdcl::FFTSO3_complex fc(L);
        const double fr1 = fr.inverse_transform(Fr, 0.4, 1.2, 2.5);
        const auto fc1 = fc.inverse_transform(Fc, 0.4, 1.2, 2.5);
        expect("real_inverse_vs_complex", std::fabs(fr1 - fc1.real()) + std::fabs(fc1.imag()), 1e-12);
    }
    return failures;
}

template <typename T> static double rel(const std::vector<T> &a, const std::vector<T> &b)
{
    double n = 0, d = 0;
    for (size_t i = 0; i < a.size(); i++) {
        n += std::norm(a[i] - b[i]);
        d += std::norm(b[i]);
    }
    return std::sqrt(n / d);
}

static int gpu_checks()
{
    std::srand(11);
    // the reference's test: round trips at l_max = 5 below 1e-10 (test/test.cpp:39-42,72-88)
    {
        fdcl::FFTSO3_complex fc(5);
        expect("so3_complex_check_transform_l5", fc.check_transform(), 1e-10);
        fdcl::FFTSO3_real fr(5);
        expect("so3_real_check_transform_l5", fr.check_transform(), 1e-10);
        fdcl::FFTS2_complex sc(5);
        expect("s2_complex_check_transform_l5", sc.check_transform(), 1e-10);
        fdcl::FFTS2_real sr(5);
        expect("s2_real_check_transform_l5", sr.check_transform(), 1e-10);
    }
    // grid inverse (GPU) against pointwise inverse (host) on a sample of grid points, then the
    // forward of the grid back to F
    {
        const int L = 12, N = 2 * (L + 1);
        fdcl::FFTSO3_complex fc(L);
        fdcl::FFTSO3_matrix_complex F(L);
        F.setRandom();
        const auto g = fc.inverse_transform(F);
        double err = 0, nrm = 0;
        for (int s = 0; s < 40; s++) {
            const int k = (7 * s) % N, j1 = (13 * s + 1) % N, j2 = (5 * s + 3) % N;
            const auto p = fc.inverse_transform(F, fc.alpha_j(j1), fc.beta_k(k), fc.gamma_j(j2));
            err += std::norm(p - g[((size_t)k * N + j1) * N + j2]);
            nrm += std::norm(p);
        }
        expect("so3_grid_inverse_vs_pointwise", std::sqrt(err / nrm), 1e-12);
        const auto F2 = fc.forward_transform(g);
        expect("so3_grid_round_trip", rel(F2.v, F.v), 1e-12);
        // batched entry points and the FP32 device path
        std::vector<fdcl::FFTSO3_matrix_complex> Fs(3, F);
        Fs[1] = F * 2.0;
        Fs[2] = F * fdcl::I;
        const auto gb = fc.inverse_transform_batch(Fs);
        const auto Fb = fc.forward_transform_batch(gb, 3);
        double eb = 0;
        for (int i = 0; i < 3; i++) eb = std::max(eb, rel(Fb[i].v, Fs[i].v));
        expect("so3_batch_round_trip", eb, 1e-12);
        fc.set_precision(FFTSO3_FP32);
        const auto F3 = fc.forward_transform(fc.inverse_transform(F));
        expect("so3_fp32_round_trip", rel(F3.v, F.v), 2e-6);
    }
    // copies of one object used from two threads at once: each copy has its own device plan, so
    // both give the result of the original (run alone) with nothing shared between them
    {
        const int L = 14;
        fdcl::FFTSO3_real proto(L);
        fdcl::FFTSO3_matrix_real F1(L), F2(L);
        F1.setRandom();
        F2.setRandom();
        const auto ref1 = proto.forward_transform(proto.inverse_transform(F1));
        const auto ref2 = proto.forward_transform(proto.inverse_transform(F2));
        std::vector<fdcl::FFTSO3_real> copies(2, proto);
        std::vector<double> errs(2, 1.0);
        std::vector<std::thread> th;
        for (int t = 0; t < 2; t++)
            th.emplace_back([&, t]() {
                try {
                    double e = 0;
                    for (int r = 0; r < 6; r++) {
                        const auto &F = (r + t) % 2 ? F2 : F1;
                        const auto &ref = (r + t) % 2 ? ref2 : ref1;
                        const auto out = copies[t].forward_transform(copies[t].inverse_transform(F));
                        e = std::max(e, rel(out.v, ref.v));
                    }
                    errs[t] = e;
                } catch (const std::exception &ex) {
                    std::printf("FAIL thread %d: %s\n", t, ex.what());
                    errs[t] = 1.0;
                }
            });
        for (auto &x : th) x.join();
        expect("copies_from_two_threads", std::max(errs[0], errs[1]), 1e-15);
    }
    // batched pointwise evaluation on the GPU against the host pointwise inverse
    {
        const int L = 10;
        fdcl::FFTSO3_complex fc(L);
        fdcl::FFTSO3_matrix_complex F(L);
        F.setRandom();
        std::vector<double> ang;
        for (int s = 0; s < 25; s++) {
            ang.push_back(0.37 * s - 2.0);
            ang.push_back(0.13 * s);
            ang.push_back(1.0 - 0.29 * s);
        }
        const auto v = fc.inverse_transform_points(F, ang);
        std::vector<fdcl::complex<double>> ref;
        for (int s = 0; s < 25; s++) ref.push_back(fc.inverse_transform(F, ang[3 * s], ang[3 * s + 1], ang[3 * s + 2]));
        expect("so3_points_vs_pointwise", rel(v, ref), 1e-12);
        fdcl::FFTSO3_real fr(L);
        fdcl::FFTSO3_matrix_real Fr(L);
        Fr.setRandom();
        const auto vr = fr.inverse_transform_points(Fr, ang);
        double e = 0, nn = 0;
        for (int s = 0; s < 25; s++) {
            const double r = fr.inverse_transform(Fr, ang[3 * s], ang[3 * s + 1], ang[3 * s + 2]);
            e += (vr[s] - r) * (vr[s] - r);
            nn += r * r;
        }
        expect("so3_real_points_vs_pointwise", std::sqrt(e / nn), 1e-12);
        fdcl::FFTS2_complex sc(L);
        fdcl::FFTS2_matrix_complex G(L);
        G.setRandom();
        std::vector<double> a2;
        for (int s = 0; s < 25; s++) {
            a2.push_back(0.12 * s);
            a2.push_back(0.4 * s - 3.0);
        }
        const auto vs = sc.inverse_transform_points(G, a2);
        std::vector<fdcl::complex<double>> rs;
        for (int s = 0; s < 25; s++) rs.push_back(sc.inverse_transform(G, a2[2 * s], a2[2 * s + 1]));
        expect("s2_points_vs_pointwise", rel(vs, rs), 1e-12);
    }
    {
        const int L = 20, N = 2 * (L + 1);
        fdcl::FFTS2_complex sc(L);
        fdcl::FFTS2_matrix_complex F(L);
        F.setRandom();
        const auto g = sc.inverse_transform(F);
        double err = 0, nrm = 0;
        for (int s = 0; s < 40; s++) {
            const int k = (7 * s) % N, j = (13 * s + 1) % N;
            const auto p = sc.inverse_transform(F, sc.theta_k(k), sc.phi_j(j));
            err += std::norm(p - g[(size_t)k * N + j]);
            nrm += std::norm(p);
        }
        expect("s2_grid_inverse_vs_pointwise", std::sqrt(err / nrm), 1e-12);
        expect("s2_grid_round_trip", rel(sc.forward_transform(g).v, F.v), 1e-12);
    }
    // S^2 matching (example/example3.cpp): the correlation grid from one GPU inverse equals the cost
    // J(R) = sum_l G^T U(R) F at grid rotations, and peaks at the rotation applied to F
    {
        const int L = 16, N = 2 * (L + 1);
        fdcl::FFTSO3_real fr(L);
        fdcl::FFTS2_matrix_real F(L);
        F.setRandom();
        const int k0 = 9, a0 = 5, g0 = 21;
        const fdcl::Matrix3d Rt = fdcl::Euler3232R(fr.alpha_j(a0), fr.beta_k(k0), fr.gamma_j(g0));
        const fdcl::FFTS2_matrix_real G = fr.rotate(F, Rt);
        const std::vector<double> J = fr.correlation(F, G);
        double err = 0, nrm = 0;
        for (int s = 0; s < 20; s++) {
            const int k = (7 * s + 2) % N, j1 = (11 * s + 1) % N, j2 = (5 * s + 3) % N;
            const fdcl::FFTSO3_matrix_real U = fr.real_harmonics(fr.alpha_j(j1), fr.beta_k(k), fr.gamma_j(j2));
            double ref = 0;
            for (int l = 0; l <= L; l++)
                for (int m = -l; m <= l; m++)
                    for (int n = -l; n <= l; n++) ref += G(l, m) * U(l, m, n) * F(l, n);
            err += std::pow(J[((size_t)k * N + j1) * N + j2] - ref, 2);
            nrm += ref * ref;
        }
        expect("s2_correlation_vs_direct_cost", std::sqrt(err / nrm), 1e-10);
        const size_t best = std::max_element(J.begin(), J.end()) - J.begin();
        expect("s2_correlation_peak_at_true_rotation", (double)(best != ((size_t)k0 * N + a0) * N + g0), 0);
    }
    // batched GPU rotation of S^2 coefficients equals the host rotation G = U(R) F (real) and
    // D(R) F computed from wigner_D (complex)
    {
        const int L = 9;
        fdcl::FFTSO3_real fr(L);
        fdcl::FFTS2_matrix_real F(L);
        F.setRandom();
        std::vector<fdcl::Matrix3d> Rs;
        for (int i = 0; i < 7; i++) Rs.push_back(fdcl::Euler3232R(0.3 + 0.7 * i, 0.2 + 0.4 * i, -1.0 + 0.5 * i));
        const auto Gb = fr.rotate_batch(F, Rs);
        double e = 0;
        for (size_t i = 0; i < Rs.size(); i++) e = std::max(e, rel(Gb[i].v, fr.rotate(F, Rs[i]).v));
        expect("s2_rotate_batch_real_vs_host", e, 1e-12);
        fdcl::FFTSO3_complex fc(L);
        fdcl::FFTS2_matrix_complex Fc(L);
        Fc.setRandom();
        const auto Gc = fc.rotate_batch(Fc, Rs);
        double ec = 0;
        for (size_t i = 0; i < Rs.size(); i++) {
            const auto D = fc.wigner_D(Rs[i]);
            std::vector<fdcl::complex<double>> ref(Fc.v.size());
            for (int l = 0; l <= L; l++)
                for (int m = -l; m <= l; m++) {
                    fdcl::complex<double> s = 0;
                    for (int n = -l; n <= l; n++) s += D(l, m, n) * Fc(l, n);
                    ref[(size_t)l * l + m + l] = s;
                }
            ec = std::max(ec, rel(Gc[i].v, ref));
        }
        expect("s2_rotate_batch_complex_vs_wigner_D", ec, 1e-12);
    }
    // forward_transform(std::function): the chunked, pipelined host sampling must give exactly the
    // coefficients of the grid overload on the same samples (l_max = 40: several 4 MB chunks)
    {
        const int L = 40, N = 2 * (L + 1);
        fdcl::FFTSO3_complex fc(L);
        auto f = [](double a, double b, double g) {
            const fdcl::Matrix3d R = fdcl::Euler3232R(a, b, g);
            return fdcl::complex<double>(R(0, 0) + R(1, 1) + R(2, 2), 0.5 * R(0, 2) - R(2, 1));
        };
        std::vector<fdcl::complex<double>> g((size_t)N * N * N);
        for (int k = 0; k < N; k++)
            for (int j1 = 0; j1 < N; j1++)
                for (int j2 = 0; j2 < N; j2++)
                    g[((size_t)k * N + j1) * N + j2] = f(fc.alpha_j(j1), fc.beta_k(k), fc.gamma_j(j2));
        const auto Fg = fc.forward_transform(g);
        const auto Fs = fc.forward_transform(std::function<fdcl::complex<double>(double, double, double)>(f));
        expect("so3_sampled_vs_grid_forward", rel(Fs.v, Fg.v), 1e-15);
        // tr R = sum_m D^1_mm: only degree 1 is nonzero, F^1 = I / 3 up to the (1,3) term's entries
        expect("so3_sampled_trR_F1_diag", std::abs(Fs(1, 0, 0) - 1.0 / 3.0), 1e-12);
        fc.set_precision(FFTSO3_FP32);
        const auto F32 = fc.forward_transform(std::function<fdcl::complex<double>(double, double, double)>(f));
        expect("so3_sampled_fp32_vs_fp64", rel(F32.v, Fg.v), 2e-6);
        fdcl::FFTSO3_real fr(L);
        auto fre = [](double a, double b, double g) {
            const fdcl::Matrix3d R = fdcl::Euler3232R(a, b, g);
            return R(0, 0) + R(1, 1) + R(2, 2) + 0.25 * R(1, 2);
        };
        std::vector<double> gr((size_t)N * N * N);
        for (int k = 0; k < N; k++)
            for (int j1 = 0; j1 < N; j1++)
                for (int j2 = 0; j2 < N; j2++)
                    gr[((size_t)k * N + j1) * N + j2] = fre(fr.alpha_j(j1), fr.beta_k(k), fr.gamma_j(j2));
        expect("so3_real_sampled_vs_grid_forward",
               rel(fr.forward_transform(std::function<double(double, double, double)>(fre)).v, fr.forward_transform(gr).v),
               1e-15);
        const int L2 = 200, N2 = 2 * (L2 + 1);
        fdcl::FFTS2_complex sc(L2);
        auto fs = [](double th, double ph) { return fdcl::complex<double>(std::cos(th) * std::sin(ph), std::sin(th)); };
        std::vector<fdcl::complex<double>> gs((size_t)N2 * N2);
        for (int k = 0; k < N2; k++)
            for (int j = 0; j < N2; j++) gs[(size_t)k * N2 + j] = fs(sc.theta_k(k), sc.phi_j(j));
        expect("s2_sampled_vs_grid_forward",
               rel(sc.forward_transform(std::function<fdcl::complex<double>(double, double)>(fs)).v, sc.forward_transform(gs).v),
               1e-15);
        fdcl::FFTS2_real sr(L2);
        auto fsr = [](double th, double ph) { return std::cos(th) + std::sin(th) * std::cos(2 * ph); };
        std::vector<double> gsr((size_t)N2 * N2);
        for (int k = 0; k < N2; k++)
            for (int j = 0; j < N2; j++) gsr[(size_t)k * N2 + j] = fsr(sr.theta_k(k), sr.phi_j(j));
        expect("s2_real_sampled_vs_grid_forward",
               rel(sr.forward_transform(std::function<double(double, double)>(fsr)).v, sr.forward_transform(gsr).v), 1e-15);
    }
    return failures;
}

int main(int argc, char **argv)
{
    const std::string mode = argc > 1 ? argv[1] : "cpu";
    try {
        if (mode == "cpu") return cpu_checks() ? 1 : 0;
        if (mode == "gpu") return gpu_checks() ? 1 : 0;
        if (mode == "cg" && argc > 3) {   // C^{l,m}_{l1,m1,l2,m2} for every (l, m, m1): "l m m1 m2 value"
            const int l1 = std::atoi(argv[2]), l2 = std::atoi(argv[3]);
            fdcl::Clebsch_Gordon_complex cg;
            cg.compute(l1, l2);
            for (int l = std::abs(l1 - l2); l <= l1 + l2; l++)
                for (int m = -l; m <= l; m++)
                    for (int m1 = -l1; m1 <= l1; m1++) {
                        const int m2 = m - m1;
                        if (std::abs(m2) <= l2) std::printf("%d %d %d %d %.17g\n", l, m, m1, m2, cg(l, m, l1, m1, l2, m2));
                    }
            return 0;
        }
        if (mode == "dump" && argc > 3) {
            fdcl::FFTSO3_complex fc(std::atoi(argv[3]));
            const auto d = fc.wigner_d(std::atof(argv[2]));
            for (long i = 0; i < d.size(); i++) std::printf("%.17g\n", d.data()[i]);
            return 0;
        }
        if (mode == "dumpnp" && argc > 3) {
            fdcl::FFTS2_complex sc(1);
            const auto nP = sc.nor_assoc_Legendre_poly(std::atof(argv[2]), std::atoi(argv[3]));
            for (long i = 0; i < nP.size(); i++) std::printf("%.17g\n", nP.data()[i]);
            return 0;
        }
    } catch (const std::exception &e) {
        std::printf("FAIL exception: %s\n", e.what());
        return 2;
    }
    std::fprintf(stderr, "usage: test_fdcl cpu|gpu|dump BETA L|dumpnp X L\n");
    return 2;
}
