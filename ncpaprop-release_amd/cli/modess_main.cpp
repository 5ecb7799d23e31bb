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

void configure(Options &p) {
    p.flag("help");
    p.flag("h");
    p.string("paramfile", "modess.param");
    p.flag("printparams");
    p.string("atmosfile");
    p.required("atmosfile");
    p.real("freq");
    p.required("freq");
    p.real_gt("freq", 0.0);
    p.string("atmosheaderfile", "");
    p.real("maxheight_km", 150.0);
    p.real_ge("maxheight_km", 0.1);
    p.real("zground_km", 0.0);
    p.integer("Nz_grid", 20000);
    p.int_ge("Nz_grid", 10);
    p.real("sourceheight_km", 0.0);
    p.real("receiverheight_km", 0.0);
    p.real("maxrange_km", 1000.0);
    p.real_ge("maxrange_km", 0.01);
    p.integer("Nrng_steps", 1000);
    p.int_gt("Nrng_steps", 0);
    p.string("ground_impedence_model", "rigid");
    p.one_of("ground_impedence_model", {"rigid"});
    p.string("use_attn_file", "");
    p.string("modal_starter_file", "");
    p.string("dispersion_file", "");
    p.flag("append_dispersion_file");
    p.string("filetag", "");
    p.flag("singleprop");
    p.flag("multiprop");
    p.radio("operation_mode", {"singleprop", "multiprop"});
    p.real("azimuth");
    p.real_ge("azimuth", 0.0);
    p.real_lt("azimuth", 360.0);
    p.required_if("azimuth", "singleprop");
    for (const char *k : {"azimuth_start", "azimuth_end", "azimuth_step"}) {
        p.real(k);
        p.real_ge(k, 0.0);
        p.real_le(k, 360.0);
        p.required_if(k, "multiprop");
    }
    for (const char *k : {"write_2d_tloss", "write_lossless", "write_phase_speeds", "write_speeds", "write_modes",
                          "write_atm_profile", "Lamb_wave_BC", "turnoff_WKB", "wvnum_filter"})
        p.flag(k);
    p.real("c_min");
    p.real_gt("c_min", 0.0);
    p.required_if("c_min", "wvnum_filter");
    p.real("c_max");
    p.real_gt("c_max", 0.0);
    p.required_if("c_max", "wvnum_filter");
    // dummy parameters the solver reads (modess_parameters.cpp:200-205)
    p.flag("broadband");
    p.real("f_min", 0.0);
    p.real("f_max", 0.0);
    p.real("f_step", 0.0);
}

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
    configure(p);
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
