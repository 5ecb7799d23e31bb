// ModBB: broadband normal modes, --dispersion (--method modess: effective sound speed;
// --method wmod: wide-angle high-Mach) then --propagation, on the GPU.  Options, defaults and
// validation follow src/modbb/modbb_parameters.cpp:6-260 and the program flow
// src/modbb/ModBB_main.cpp:33-121.
#include <ctime>
#include <fstream>
#include <iostream>
#include <stdexcept>

#include "ess_run.h"
#include "params.h"

namespace ncpacli {
int bb_run(const Options &p);
}

using namespace ncpacli;

namespace {

void configure(Options &p) {
    p.flag("help");
    p.flag("h");
    p.string("paramfile", "modbb.param");
    p.flag("printparams");
    p.flag("dispersion");
    p.flag("propagation");
    p.radio("operation_mode", {"dispersion", "propagation"});
    p.string("method", "modess");
    p.one_of("method", {"modess", "wmod"});
    p.required_if("method", "dispersion");
    p.string("dispersion_file", "");
    p.required_if("dispersion_file", "dispersion");
    p.flag("append_dispersion_file", true);
    p.real("azimuth");
    p.real_ge("azimuth", 0.0);
    p.real_lt("azimuth", 360.0);
    p.required_if("azimuth", "dispersion");
    for (const char *k : {"f_min", "f_max", "f_step"}) {
        p.real(k);
        p.real_gt(k, 0.0);
        p.required_if(k, "dispersion");
    }
    p.string("atmosfile");
    p.required_if("atmosfile", "dispersion");
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
    p.flag("Lamb_wave_BC");
    p.flag("write_atm_profile");
    p.string("use_attn_file", "");
    p.flag("use_zero_attenuation");
    p.flag("wvnum_filter");
    p.real("c_min");
    p.real_gt("c_min", 0.0);
    p.required_if("c_min", "wvnum_filter");
    p.real("c_max");
    p.real_gt("c_max", 0.0);
    p.required_if("c_max", "wvnum_filter");
    p.string("input_dispersion_file", "");
    p.required_if("input_dispersion_file", "propagation");
    p.string("output_waveform_file", "");
    p.required_if("output_waveform_file", "propagation");
    p.string("source", "impulse");
    p.one_of("source", {"impulse", "pulse1", "pulse2", "waveform", "spectrum"});
    p.string("source_file", "");
    p.real("f_center", -1.0);
    p.integer("nfft", 0);
    p.int_ge("nfft", 0);
    p.real("max_celerity", 340.0);
    p.int_gt("max_celerity", 0);
    p.string("receiver", "single");
    p.one_of("receiver", {"single", "multiple"});
    p.real("range_km", 0.0);
    p.real("start_range_km", 0.0);
    p.real("end_range_km", 0.0);
    p.real("range_step_km", 0.0);
    p.flag("broadband", true);
    // dummy parameters the solver reads (modbb_parameters.cpp:223-234)
    p.string("modal_starter_file", "");
    for (const char *k : {"write_2d_tloss", "write_phase_speeds", "write_speeds", "write_modes", "write_lossless",
                          "multiprop", "turnoff_WKB"})
        p.flag(k);
    p.string("filetag", "");
}

void usage(std::ostream &os) {
    os << "ModBB -- broadband normal modes (effective sound speed, MI355X)\n\n"
          "1. --dispersion --dispersion_file <file> --atmosfile <file> --azimuth <deg>\n"
          "   --f_min <Hz> --f_step <Hz> --f_max <Hz> [--method modess]\n"
          "2. --propagation --input_dispersion_file <file> --output_waveform_file <file>\n"
          "   --source {impulse,pulse1,pulse2,spectrum,waveform} [--source_file <file>] [--f_center <Hz>]\n"
          "   --receiver single --range_km <km> | --receiver multiple --start_range_km --end_range_km\n"
          "   --range_step_km; [--nfft N] [--max_celerity 340] [--use_zero_attenuation]\n"
          "Optional (dispersion): --atmosheaderfile --maxheight_km --zground_km --Nz_grid\n"
          "   --sourceheight_km --receiverheight_km --maxrange_km --Nrng_steps --Lamb_wave_BC\n"
          "   --use_attn_file --wvnum_filter --c_min --c_max --paramfile [modbb.param] --printparams\n";
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
        return 0;
    }
    if (!p.validate()) {
        std::cout << "Parameter validation failed:" << std::endl;
        p.print_failed(std::cout);
        return 0;
    }
    const std::time_t t1 = std::time(nullptr);
    int rc = 0;
    try {
        if (p.get_bool("dispersion")) {
            const std::string disp = p.get_string("dispersion_file");
            if (!disp.empty()) std::ofstream(disp, std::ios_base::trunc).close();   // ModBB_main.cpp:77-82
            rc = ess_run(ess_from_options(p));
            if (rc == 0) std::cout << "Finished computing and saving dispersion data" << std::endl;
        } else if (p.get_bool("propagation")) {
            rc = bb_run(p);
        }
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    std::cout << "\nRun duration: " << std::difftime(std::time(nullptr), t1) << " seconds." << std::endl;
    std::cout << " ... main() broadband version is done." << std::endl;
    return rc;
}
