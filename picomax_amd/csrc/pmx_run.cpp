/*
 * pmx_run.cpp -- the body of the `picomax` CLI after option parsing: the
 * reference's main (src/pmx.cpp:43-207) minus the set-up that pmx_input.cpp
 * does, with the same stdout progress lines and the same output files:
 *   band   -> ${wdir}/${outputfile}.dat  [NQ][NBAND] f64   (export_eband,
 *             src/pmx_io.cpp:704-713)
 *   epsij  -> ${wdir}/${outputfile}.dat  Re then Im, [i][j][q][m][n<=m][f]
 *             (export_Xij, src/pmx_io.cpp:720-741) and .gpt [NEPS][3] f64
 *             (export_gpt, src/pmx_io.cpp:816-826)
 *   epsLL  -> the same two files from chi_LL, dispatched as in the older
 *             snapshot's main (src_bkp/pmx.cpp:206-224: setRefEnergy, chi_LL,
 *             export_Xij, export_gpt); the current main has no such branch
 *             (src/pmx.cpp:171-203)
 *   epsxyz -> the same two files from chi_xyz (src_bkp/pmx.cpp:226-244)
 */
#include "pmx_internal.h"

#include <chrono>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

namespace {

using clk = std::chrono::system_clock;

double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

int write_doubles(const std::string &path, const std::vector<const double *> &parts, size_t n_each) {
    std::ofstream f(path, std::ios::binary);
    if (!f) return pmx::set_error(PMX_EIO, "cannot write " + path);
    for (const double *p : parts) f.write(reinterpret_cast<const char *>(p), sizeof(double) * n_each);
    if (!f) return pmx::set_error(PMX_EIO, "write failed: " + path);
    return PMX_OK;
}

} // namespace

extern "C" int pmx_run(pmx_ctx *ctx, const pmx_input *pin, int precision) {
    if (!pin) return pmx::set_error(PMX_EINVAL, "null input");
    const pmx::Input &in = *reinterpret_cast<const pmx::Input *>(pin);
    pmx_system sys;
    int rc = pmx_input_system(pin, &sys);
    if (rc) return rc;
    const std::string base = in.wdir + "/" + in.outputfile;
    const bool band = in.sw == "band", epsij = in.sw == "epsij", epsll = in.sw == "epsLL",
               epsxyz = in.sw == "epsxyz";
    if (!band && !epsij && !epsll && !epsxyz) return PMX_OK; /* unknown or absent switch: nothing, exit 0 */
    if (!ctx) return pmx::set_error(PMX_EINVAL, "null context");
    if ((rc = pmx_set_system(ctx, &sys))) return rc;

    if (band) {
        std::cout << "Electronic bandstructure solver" << std::endl;
        auto t0 = clk::now();
        double off = 0.0;
        const double rp[3] = {in.refpoint.x, in.refpoint.y, in.refpoint.z};
        if ((rc = pmx_ref_energy(ctx, rp, in.nvalence, &off))) return rc;
        if (!(in.nvalence > 0 && in.nvalence <= 2 * in.npw))
            std::cout << "WARNING::setRefEnergy:: nvalence is out of bounds for eigenvalues. Setting offset to 0."
                      << std::endl;
        const int nq = (int)in.Q.size();
        std::vector<double> eb((size_t)nq * in.nband);
        if ((rc = pmx_bands(ctx, nq, in.Q_flat.data(), in.nband, off, eb.data()))) return rc;
        std::cout << "  Elapsed time: " << secs(t0, clk::now()) << " s" << std::endl;
        return write_doubles(base + ".dat", {eb.data()}, eb.size());
    }

    /* epsij / epsLL / epsxyz */
    std::cout << (epsxyz ? "Susceptibility tensor matrix solver (xyz-basis)" : "Susceptibility tensor matrix solver")
              << std::endl;
    auto t0 = clk::now();
    if (!in.has_freq) return pmx::set_error(PMX_EINVAL, "-switch " + in.sw + " needs both -nfreq and -dfreq");
    double off = 0.0;
    const double rp[3] = {in.refpoint.x, in.refpoint.y, in.refpoint.z};
    if ((rc = pmx_ref_energy(ctx, rp, in.nvalence, &off))) return rc;
    pmx_chi_params p;
    std::memset(&p, 0, sizeof(p));
    p.nk = (int)in.K.size();
    p.K = in.K_flat.data();
    p.KW = in.KW.data();
    p.nq = (int)in.Q.size();
    p.Q = in.Q_flat.data();
    p.nband = in.nband;
    p.offset = off;
    p.delta = in.delta;
    p.epsilon = in.epsilon;
    p.nfreq = in.nfreq;
    p.dfreq = in.dfreq;
    p.kk = in.kk;
    p.precision = precision;
    const size_t ne2 = (size_t)in.neps * (in.neps + 1) / 2;
    const size_t n = 9 * (size_t)p.nq * ne2 * in.nfreq;
    std::vector<double> re(n), im(n);
    pmx::set_progress(ctx, true);
    rc = (epsxyz ? pmx_chi_xyz : epsll ? pmx_chi_ll : pmx_chi)(ctx, &p, re.data(), im.data());
    pmx::set_progress(ctx, false);
    if (rc) return rc;
    if ((rc = write_doubles(base + ".dat", {re.data(), im.data()}, n))) return rc;
    if ((rc = write_doubles(base + ".gpt", {in.G.data()}, (size_t)3 * in.neps))) return rc;
    std::cout << "  Elapsed time: " << secs(t0, clk::now()) << " s" << std::endl;
    return PMX_OK;
}
