// The normal-mode run behind the Modess, WMod and ModBB --dispersion command lines:
// ESSModeSolver / WModeSolver setParams + solve (src/libmodes/ESSModeSolver.cpp,
// WModeSolver.cpp) on the GPU through the ncpamodes C ABI, with the reference's output files
// and formats.
#pragma once
#include <string>
#include <vector>

#include "params.h"

namespace ncpacli {

struct EssRun {
    std::string atmosfile, atmosheaderfile, usrattfile, modstartfile, dispersion_file, user_tag;
    std::string gnd_imp_model = "rigid";
    double z_min = 0.0, maxrange = 1000e3, maxheight = 150e3, sourceheight = 0.0, receiverheight = 0.0;
    int Nz_grid = 20000, Nrng_steps = 1000;
    bool Lamb_wave_BC = false, write_2D_TLoss = false, write_atmosphere = false, write_phase_speeds = false;
    bool write_speeds = false, write_modes = false, append_dispersion_file = false, Nby2Dprop = false;
    bool turnoff_WKB = false, write_lossless = false, wvnum_filter = false;
    double c_min = 0.0, c_max = 0.0;
    std::vector<double> freqs;     // f_vec (one entry unless broadband)
    double azi_min = 0.0, azi_step = 0.0;
    int Naz = 1;
    bool wide = false;             // WModeSolver (WMod, ModBB --method wmod) instead of ESSModeSolver
};

// Options of Modess, or (wide) of WMod
void configure_modes_options(Options &p, bool wide);

// ESSModeSolver::setParams option reading (ESSModeSolver.cpp:30-124; WModeSolver::setParams,
// WModeSolver.cpp:30-115, reads the same options)
EssRun ess_from_options(const Options &p);

// Runs every (frequency, azimuth) problem as one batch on the GPU and writes the outputs.
// Returns 0 on success; prints the error and returns non-zero otherwise.
int ess_run(const EssRun &r);

// Throws std::runtime_error("cannot open <path> ...") unless the run can write `path`: its
// directory is writable and so is the file if it exists.
void check_writable(const std::string &path);

}  // namespace ncpacli
