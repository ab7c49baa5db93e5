// Host check of the kernels' pair evaluation: the owner-hoisted, merged-regime form used by the
// bath-rate and chain kernels (make_owner + catastrophic_moments + eval_pair_terms) against the
// direct per-pair forms (vrel_pair + outcome_moments, one expression per reference line), bit for
// bit, over random owners and partners that cover every turbulence regime and outcome branch.
//
//   pair_forms_check <trials> <seed>       exit 0 when every field of every trial is identical
//   pair_forms_check log <trials> <seed>   ulog1p / ulog against extended-precision log1pl / logl:
//                                          exit 0 when every error is below one ulp
//
// Built for the host only (hipcc -x hip --cuda-host-only) from the same header the kernels use.
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "gdx_device.h"

using namespace gdx;

namespace {

struct lcg
{
    uint64_t s;
    double next()
    {
        s = s*6364136223846793005ull + 1442695040888963407ull;
        return (double((s >> 11)) + 0.5)*0x1p-53;
    }
};

uint64_t bits(double v)
{
    uint64_t b;
    std::memcpy(&b, &v, sizeof b);
    return b;
}

// the reference's per-pair statement order (col_bath_rate.cu:40-57, _col_rates.cuh:127-141)
pair_term direct(const gdx_env &e, double size_i, double m, double hg_i, double size_j, double numr_j,
                 double hg_j, double lambda_0)
{
    pair_term p{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double vrel = 0.0;
    double rate;
    if constexpr (COAG_KERNEL == 0) rate = lambda_0*numr_j;
    else if constexpr (COAG_KERNEL == 1) rate = lambda_0*numr_j*(grain_mass(size_i) + grain_mass(size_j));
    else if constexpr (COAG_KERNEL == 2) rate = lambda_0*numr_j*grain_mass(size_i)*grain_mass(size_j);
    else
    {
        vrel = vrel_pair(e, size_i, size_j);
        rate = numr_j*vrel*PI*(size_i + size_j)*(size_i + size_j) / 4.0;
        if constexpr (N_Z == 1) rate /= sqrt(2.0*PI*(hg_i*hg_i + hg_j*hg_j));
    }
    double r = rate / m;
    double mean = 0.0, second = 0.0, maximum = 0.0;
    r *= outcome_moments(size_i, size_j, vrel >= V_FRAG, mean, second, maximum);
    p.rate = r;
    p.j1 = r*mean;
    p.j2 = r*second;
    p.jmax = r > 0.0 ? maximum : 0.0;
    p.vrel = vrel;
    p.size_j = size_j;
    return p;
}

int regime(const gdx_env &e, double size_i, double size_j)
{
    double si = stokes(e, size_i), sj = stokes(e, size_j);
    double sl = si > sj ? si : sj, ris = e.re_inv_sqrt;
    return sl < 0.2*ris ? 0 : sl < ris / 1.6 ? 1 : sl < 5.0*ris ? 2 : sl < 0.2 ? 3 : sl < 1.0 ? 4 : 5;
}

// error of y in units of the last place of the double nearest to the exact value t
double ulp_error(double y, long double t)
{
    double tr = static_cast<double>(t);
    if (tr == 0.0) return y == 0.0 ? 0.0 : 1e30;
    int e;
    std::frexp(tr, &e);
    double ulp = std::ldexp(1.0, e - 53);
    return static_cast<double>(std::fabs(static_cast<long double>(y) - t) / ulp);
}

int check_logs(long long trials, lcg &g)
{
    double worst1 = 0.0, worst2 = 0.0, x1 = 0.0, x2 = 0.0;
    long long differ1 = 0, differ2 = 0;
    for (long long n = 0; n < trials; ++n)
    {
        // log1p arguments: tiny to huge magnitudes of both signs above -1, and the reduction edges
        double x;
        double r = g.next();
        if (r < 0.3) x = std::ldexp(g.next(), -int(60*g.next()));
        else if (r < 0.5) x = -std::ldexp(g.next(), -int(60*g.next()));
        else if (r < 0.7) x = -0.3 + 0.75*g.next();
        else if (r < 0.8) x = std::sqrt(0.5) - 1.0 + (g.next() - 0.5)*1e-6;
        else if (r < 0.9) x = std::sqrt(2.0) - 1.0 + (g.next() - 0.5)*1e-6;
        else x = std::pow(10.0, 300.0*g.next());
        double y = ulog1p(x);
        double err = ulp_error(y, log1pl(static_cast<long double>(x)));
        if (err > worst1) { worst1 = err; x1 = x; }
        differ1 += y != std::log1p(x);
        // log arguments: mass ratios over the whole normal range
        double q = std::pow(10.0, -300.0 + 600.0*g.next());
        double yq = ulog(q);
        double eq = ulp_error(yq, logl(static_cast<long double>(q)));
        if (eq > worst2) { worst2 = eq; x2 = q; }
        differ2 += yq != std::log(q);
    }
    bool special = ulog1p(0.0) == 0.0 && std::isinf(ulog1p(-1.0)) && ulog1p(-1.0) < 0 && std::isnan(ulog1p(-2.0))
        && std::isnan(ulog1p(__builtin_nan(""))) && std::isinf(ulog1p(__builtin_inf())) && ulog1p(1e-300) == 1e-300
        && ulog(1.0) == 0.0 && std::isinf(ulog(0.0)) && std::isnan(ulog(-1.0)) && std::isinf(ulog(__builtin_inf()));
    std::printf("log1p worst_ulp %.4f at %.17g differs_from_libm %lld\n", worst1, x1, differ1);
    std::printf("log worst_ulp %.4f at %.17g differs_from_libm %lld\n", worst2, x2, differ2);
    std::printf("special %d\n", int(special));
    return (worst1 < 1.0 && worst2 < 1.0 && special) ? 0 : 1;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc > 1 && std::strcmp(argv[1], "log") == 0)
    {
        lcg lg{argc > 3 ? std::strtoull(argv[3], nullptr, 10) : 1ull};
        return check_logs(argc > 2 ? std::atoll(argv[2]) : 1000000, lg);
    }
    long long trials = argc > 1 ? std::atoll(argv[1]) : 1000000;
    lcg g{argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 1ull};
    long long bad[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, reg[6] = {0, 0, 0, 0, 0, 0}, branch[3] = {0, 0, 0};
    for (long long n = 0; n < trials; ++n)
    {
        double y = Y_MIN + (Y_MAX - Y_MIN)*g.next();
        double z = (N_Z > 1) ? Z_MIN + (Z_MAX - Z_MIN)*g.next() : 0.5*PI;
        gdx_env e = make_env(y, z);
        // sizes log-uniform over eight decades around the reference size, so the larger Stokes
        // number visits every turbulence regime and the mass ratio every outcome branch
        double size_i = S_0*std::pow(10.0, -4.0 + 8.0*g.next());
        double size_j = (g.next() < 0.25) ? size_i*std::pow(10.0, -0.5 + g.next()) : S_0*std::pow(10.0, -4.0 + 8.0*g.next());
        double numr_j = std::pow(10.0, -3.0 + 6.0*g.next());
        double m = std::pow(10.0, -6.0 + 6.0*g.next());
        double R = cyl_R(y, z);
        double hg_i = hg(R), hg_j = hg(R*(0.9 + 0.2*g.next()));
        double lambda_0 = 1.0 + g.next();

        pair_term a = direct(e, size_i, m, hg_i, size_j, numr_j, hg_j, lambda_0);
        owner_terms t = make_owner(e, size_i, m, hg_i);
        cat_moments cm = catastrophic_moments(size_i);
        partner q{size_j, numr_j, hg_j, true};
        pair_term b = eval_pair_terms(t, cm, q, lambda_0);
        // the rate-only form used for owners without compensators gives the same rate and speed
        pair_term c = eval_pair_rate(t, q, lambda_0);
        if ((bits(c.rate) != bits(b.rate) && !(std::isnan(c.rate) && std::isnan(b.rate))) || bits(c.vrel) != bits(b.vrel))
        {
            if (bad[5] < 3) std::printf("rate-only mismatch: size_i %.17g size_j %.17g\n", size_i, size_j);
            bad[5]++;
        }

        // the chain's held partner factors, taken at the bath start (another owner size) and kept across
        // the owner's resize after an event: the same terms bit for bit
        owner_terms t0 = make_owner(e, size_i*(0.5 + g.next()), m, hg_i);
        const partner_terms held = make_partner_terms(t0, size_j);
        owner_resize(t0, size_i);
        pair_term h = eval_pair_terms<cat_moments, true>(t0, cm, q, lambda_0, held);
        pair_term hr = eval_pair_rate<true>(t0, q, lambda_0, held);
        const double hv[6] = {h.rate, h.j1, h.j2, h.jmax, h.vrel, hr.rate}, bw[6] = {b.rate, b.j1, b.j2, b.jmax, b.vrel, b.rate};
        for (int f = 0; f < 6; ++f)
            if (bits(hv[f]) != bits(bw[f]) && !(std::isnan(hv[f]) && std::isnan(bw[f])))
            {
                if (bad[6] < 3) std::printf("held-partner mismatch field %d: size_i %.17g size_j %.17g\n", f, size_i, size_j);
                bad[6]++;
            }

        // the chain's deferred catastrophic moments: zeros in the evaluation, the moments filled in for
        // the pairs that took the catastrophic branch
        pair_term d = eval_pair_terms(t, cat_moments_deferred{}, q, lambda_0);
        if (catastrophic_pair(t, size_j, d.vrel >= V_FRAG))
        {
            d.j1 = d.rate*cm.mean;
            d.j2 = d.rate*cm.second;
            d.jmax = d.rate > 0.0 ? cm.maximum : 0.0;
        }
        const double dv[5] = {d.rate, d.j1, d.j2, d.jmax, d.vrel}, dw[5] = {b.rate, b.j1, b.j2, b.jmax, b.vrel};
        for (int f = 0; f < 5; ++f)
            if (bits(dv[f]) != bits(dw[f]) && !(std::isnan(dv[f]) && std::isnan(dw[f])))
            {
                if (bad[7] < 3) std::printf("deferred-moment mismatch field %d: size_i %.17g size_j %.17g\n", f, size_i, size_j);
                bad[7]++;
            }
        // the chain's outcome through the owner's held reciprocal: the same new size
        const double u = g.next();
        for (int hs = 0; hs < 2; ++hs)
        {
            const double s1 = sample_outcome(size_i, size_j, hs == 1, u), s2 = sample_outcome_owner(t, size_j, hs == 1, u);
            if (bits(s1) != bits(s2) && !(std::isnan(s1) && std::isnan(s2)))
            {
                if (bad[8] < 3) std::printf("outcome mismatch: size_i %.17g size_j %.17g u %.17g\n", size_i, size_j, u);
                bad[8]++;
            }
        }

        if constexpr (COAG_KERNEL == 3) reg[regime(e, size_i, size_j)]++;
        double ratio = size_j / size_i;
        bool high = a.vrel >= V_FRAG;
        branch[!high ? 0 : (ratio*ratio*ratio <= 0.1 ? 1 : 2)]++;
        const double av[5] = {a.rate, a.j1, a.j2, a.jmax, a.vrel}, bv[5] = {b.rate, b.j1, b.j2, b.jmax, b.vrel};
        for (int f = 0; f < 5; ++f)
        {
            if (bits(av[f]) != bits(bv[f]) && !(std::isnan(av[f]) && std::isnan(bv[f])))
            {
                if (bad[f] < 3)
                    std::printf("mismatch field %d: size_i %.17g size_j %.17g y %.17g z %.17g: %.17g vs %.17g\n", f,
                                size_i, size_j, y, z, av[f], bv[f]);
                bad[f]++;
            }
        }
    }
    std::printf("model %s trials %lld\n", GDX_MODEL_NAME, trials);
    std::printf("regimes %lld %lld %lld %lld %lld %lld\n", reg[0], reg[1], reg[2], reg[3], reg[4], reg[5]);
    std::printf("branches %lld %lld %lld\n", branch[0], branch[1], branch[2]);
    std::printf("mismatches %lld %lld %lld %lld %lld %lld %lld %lld %lld\n", bad[0], bad[1], bad[2], bad[3], bad[4],
                bad[5], bad[6], bad[7], bad[8]);
    return (bad[0] | bad[1] | bad[2] | bad[3] | bad[4] | bad[5] | bad[6] | bad[7] | bad[8]) ? 1 : 0;
}
