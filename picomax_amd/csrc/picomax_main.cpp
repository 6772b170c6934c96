/*
 * picomax_main.cpp -- the `picomax` command line, a drop-in for the
 * reference's CLI (src/pmx.cpp:43-207): same options, same progress lines,
 * same .dat/.gpt files, same exit behaviour (abort() where the reference
 * aborts, 0 on success and on an unknown or absent -switch).  All computation
 * goes through libpmx (include/pmx.h) on GPU 0 (or $PMX_DEVICE).
 *
 * One option is added: -precision fp64|fp32|bf16x3|direct selects the chi path
 * (default fp64, the plain path).  The reference ignores unknown options, so
 * command lines stay interchangeable.
 */
#include "pmx.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

static void doc() {
    std::cout << "============================================================\n"
              << "  PicoMax (MI355X) version 0.1\n"
              << "Usage:\n"
              << "  picomax [-option] [value]\n"
              << "============================================================\n"
              << "-h, -help : print help\n"
              << "-v, -version : print PicoMax version\n"
              << "-debug : debug level {0}\n"
              << "-outputfile : output filename\n"
              << "-wdir : working directory {pwd}\n"
              << "-logfile : print output to a logfile {0}\n\n"
              << "-switch : calculation switch\n"
              << "    band, epsij, epsLL\n\n"
              << "-crystal : crystal structure\n"
              << "    zb, wz, rs, te, hex\n"
              << "-dim : crystal dimension {3}\n"
              << "-a : lattice constant [angstrom]\n"
              << "-f : lattice parameter\n"
              << "-u : atomic basis parameter\n"
              << "-vg : empirical pseudopotential form factors\n"
              << "-us_so : Symmetric Spin-Orbit form factors (G^2:value,G^2:value,...)\n"
              << "-use_uniform_us_so : Use a single value for symmetric SO coupling for all G^2 (flag)\n"
              << "-uniform_us_so_val : The uniform symmetric SO coupling value in Rydbergs\n"
              << "-ua_so : Antisymmetric Spin-Orbit form factors (G^2:value,G^2:value,...)\n"
              << "-ua_so_val_g0 : Antisymmetric Spin-Orbit form factor for G^2=0 (single value)\n"
              << "-encut : energy cutoff [eV]\n"
              << "-gcut : g-vector cutoff [2*pi/a]\n"
              << "-gsym : spherical (1) or rectangular (0) g-vector grid {1}\n"
              << "-kpointfile : k-point file (NKPT0, NKPT, then kx ky kz w)\n"
              << "-nband : number of electronic bands\n"
              << "-neps : the size of dielectric matrix NxN\n"
              << "-nvalence : the number of valence electron bands\n\n"
              << "-qvec : q-vectors\n"
              << "-qpath : path of q-vectors\n"
              << "-qnum : number of q-vectors in each path\n"
              << "-refpoint : the reference q-vector for electronic band energy {0,0,0}\n"
              << "-gammazero : the Gamma-point {0,0,0}\n\n"
              << "-nfreq : number of frequency points {1}\n"
              << "-dfreq : frequency interval [eV]\n"
              << "-delta : delta function model {0}: 0 gaussian, 1 lorentzian, 2 triangle, 3 rect\n"
              << "-epsilon : delta function smearing [eV] {0.1}\n"
              << "-kk : switch for Kramers-Kronig transform {0}\n"
              << "-precision : chi path fp64 (default) | fp32 | bf16x3 | direct\n"
              << "============================================================\n"
              << std::endl;
}

static bool has(int argc, char **argv, const char *o) {
    for (int i = 1; i < argc; i++)
        if (!std::strcmp(argv[i], o)) return true;
    return false;
}

static const char *value(int argc, char **argv, const char *o) {
    for (int i = 1; i < argc; i++)
        if (!std::strcmp(argv[i], o)) return i + 1 < argc ? argv[i + 1] : "";
    return "";
}

int main(int argc, char **argv) {
    if (has(argc, argv, "-h") || has(argc, argv, "-help")) {
        doc();
        return 0;
    }
    if (has(argc, argv, "-v") || has(argc, argv, "-version")) {
        std::cout << "PicoMax version v0.1 (" << pmx_version() << ")\n";
        return 0;
    }
    pmx_input *in = nullptr;
    if (pmx_input_parse(argc, argv, &in) != PMX_OK) {
        std::cout << pmx_last_error() << std::endl;
        abort();
    }
    pmx_run_info ri;
    pmx_input_info(in, &ri);
    int precision = PMX_PREC_FP64;
    const std::string prec = has(argc, argv, "-precision") ? value(argc, argv, "-precision")
                                                           : (getenv("PMX_PRECISION") ? getenv("PMX_PRECISION") : "fp64");
    if (prec == "fp32")
        precision = PMX_PREC_FP32;
    else if (prec == "bf16x3")
        precision = PMX_PREC_BF16X3;
    else if (prec == "direct")
        precision = PMX_PREC_DIRECT;

    if (has(argc, argv, "-logfile") && std::atoi(value(argc, argv, "-logfile"))) {
        std::string lf = std::string(ri.wdir) + "/" + ri.outputfile + ".log";
        std::cout << "Output is redirected to logfile, " << lf << std::endl;
        if (!std::freopen(lf.c_str(), "w", stdout)) {
            std::cout << "Logfile may have a problem?" << std::endl;
            abort();
        }
    }
    /* banner, parameters, Q / G / K set-up lines (src/pmx.cpp:84-160) and,
     * with -debug > 1, printInput and exit before any solve (src/pmx.cpp:164-167) */
    pmx_print_setup(in, precision);
    if (ri.debug > 1) {
        pmx_input_free(in);
        return 0;
    }
    const std::string sw = ri.sw;
    pmx_ctx *ctx = nullptr;
    if (sw == "band" || sw == "epsij" || sw == "epsLL" || sw == "epsxyz") {
        const char *dev = getenv("PMX_DEVICE");
        if (pmx_create(dev ? std::atoi(dev) : 0, &ctx) != PMX_OK) {
            std::cout << "ERROR: " << pmx_last_error() << std::endl;
            abort();
        }
    }
    int rc = pmx_run(ctx, in, precision);
    if (rc != PMX_OK) {
        std::cout << "ERROR: " << pmx_last_error() << std::endl;
        abort();
    }
    pmx_destroy(ctx);
    pmx_input_free(in);
    return 0;
}
