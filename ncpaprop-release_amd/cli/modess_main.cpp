// Modess: effective-sound-speed normal modes, single azimuth or N x 2-D, on the GPU.
// Options, defaults and validation follow src/modess/modess_parameters.cpp:6-222 and the
// program flow src/modess/Modess_main.cpp:44-116.
#include <ctime>
#include <iostream>
#include <stdexcept>

#include "ess_run.h"
#include "params.h"

using namespace ncpacli;

namespace {

void usage(std::ostream &os) {
    os << "Modess -- effective sound speed normal modes (MI355X)\n\n"
          "Required: --atmosfile <file> --freq <Hz> and one of --singleprop --azimuth <deg> |\n"
          "          --multiprop --azimuth_start <deg> --azimuth_end <deg> --azimuth_step <deg>\n"
          "Optional: --atmosheaderfile --maxheight_km [150] --zground_km [0] --Nz_grid [20000]\n"
          "          --sourceheight_km [0] --receiverheight_km [0] --maxrange_km [1000] --Nrng_steps [1000]\n"
          "          --ground_impedence_model [rigid] --use_attn_file --modal_starter_file --dispersion_file\n"
          "          --append_dispersion_file --filetag --paramfile [modess.param] --printparams\n"
          "Flags:    --write_2d_tloss --write_lossless --write_phase_speeds --write_speeds --write_modes\n"
          "          --write_atm_profile --Lamb_wave_BC --turnoff_WKB --wvnum_filter --c_min --c_max\n"
          "Outputs:  tloss_1d.nm, tloss_1d.lossless.nm, tloss_2d.nm, tloss_2d.lossless.nm,\n"
          "          Nby2D_tloss_1d.nm, Nby2D_tloss_1d.lossless.nm, phasespeeds.nm, speeds.nm,\n"
          "          mode_<n>.nm, atm_profile.nm, <dispersion_file>\n";
}

}  // namespace

int main(int argc, char **argv) {
    Options p;
    configure_modes_options(p, false);
    try {
        p.parse_command_line(argc, argv);
        if (p.was_found("help") || p.was_found("h")) {
            usage(std::cout);
            return 1;
        }
        p.parse_file(p.get_string("paramfile"));
        p.parse_command_line(argc, argv);
    } catch (const std::invalid_argument &e) {
        std::cout << "Parameter validation failed: " << std::endl << e.what() << std::endl;
    }
    if (!p.validate()) {
        std::cout << "Parameter validation failed:" << std::endl;
        p.print_failed(std::cout);
        return 0;
    }
    const std::time_t t1 = std::time(nullptr);
    int rc;
    try {
        rc = ess_run(ess_from_options(p));
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    std::cout << std::endl << " ... main() is done." << std::endl;
    std::cout << "Run duration: " << std::difftime(std::time(nullptr), t1) << " seconds." << std::endl;
    return rc;
}
