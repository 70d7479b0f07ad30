// tictac_run: the reference's `run` driver over libtictac.
//
//   options, input files, defaults   CPP/set_run_parameters.cpp:76-687
//   channel and parameter-set loops  CPP/main.cpp:84-500
//   parameter file, energy file      CPP/disk_io_routines.cpp:43-110
//   U_PW_elements*.txt               CPP/disk_io_routines.cpp:759-913 (store_U_matrix_elements_txt)
//   U_PW_breakup_elements*.txt       CPP/disk_io_routines.cpp:341-565,566-706
//
//   P123 files in P123_folder        CPP/make_permutation_matrix.cpp:1297-1403 (libtictac_h5)
//
// The same key=value options (command line or <file>.txt, the last value read wins), the same
// loops and the same output files.  With calculate_and_store_P123=true P123 is built on the GPU
// (or, with P123_recovery=true, merged from the thread spill files of P123_omp_num_threads
// threads) and stored in the reference's HDF5 file; with false it is read from that file.  G is
// computed on the device inside the solve.  Differences: errors exit with status 1 (the
// reference exits with 0); the potential models are those of the library (malfliet_tjon,
// synthetic, the latter with its 15 parameters from parameter_file); one option is added,
// solve_method=pade|dense|gmres|mixed (solve_tol for the last two).
#include <sys/stat.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "tictac.h"
#include "tictac_h5.h"

namespace {

using cdouble = std::complex<double>;

[[noreturn]] void die(const std::string &msg) {
    std::cerr << "ERROR: " << msg << std::endl;
    std::exit(1);
}

void check(tt_status s) {
    if (s != TT_OK) die(tt_last_error());
}

void check_h5(tt_status s) {
    if (s != TT_OK) die(tt_h5_last_error());
}

// omp_get_max_threads() of the reference's default P123_omp_num_threads: OMP_NUM_THREADS when
// set, else the host's hardware threads
int max_threads() {
    const char *env = std::getenv("OMP_NUM_THREADS");
    const int n = env ? std::atoi(env) : (int)std::thread::hardware_concurrency();
    return n > 0 ? n : 1;
}

// run_params of CPP/type_defs.h:87-123 with the defaults of set_run_parameters.cpp:577-615
struct Options {
    int two_J_3N_max = 1, Np_WP = 50, Nq_WP = 50, J_2N_max = 3, Nphi = 48, Nx = 48;
    double chebyshev_t = 1, chebyshev_s = 100;
    int Np_per_WP = 8, Nq_per_WP = 8, channel_idx = -1, P123_omp_num_threads = max_threads();
    int PSI_start = -1, PSI_end = -1;
    bool parallel_run = false, P123_recovery = false, tensor_force = true, isospin_breaking_1S0 = true;
    bool midpoint_approx = false, calculate_and_store_P123 = true, include_breakup_channels = false;
    bool solve_faddeev = true, solve_dense = false, production_run = true, parameter_walk = false;
    std::string potential_model = "LO_internal", subfolder = "Output", p_grid_type = "chebyshev";
    std::string p_grid_filename, q_grid_type = "chebyshev", q_grid_filename, parameter_file = "none";
    std::string energy_input_file = "lab_energies.txt", output_folder = "Output", P123_folder = "Output";
    std::string solve_method = "pade";
    double solve_tol = 1e-12;
};

bool parse_bool(const std::string &option, const std::string &input) {
    if (input == "true" || input == "false") return input == "true";
    die("Invalid value for input parameter " + option + "!");
}

// read_and_set_parameter (set_run_parameters.cpp:76-251): false for an unknown option
bool set_option(Options &o, const std::string &option, const std::string &input) {
    const std::map<std::string, int *> ints = {
        {"two_J_3N_max", &o.two_J_3N_max}, {"Np_WP", &o.Np_WP}, {"Nq_WP", &o.Nq_WP},
        {"J_2N_max", &o.J_2N_max}, {"Nphi", &o.Nphi}, {"Nx", &o.Nx}, {"Np_per_WP", &o.Np_per_WP},
        {"Nq_per_WP", &o.Nq_per_WP}, {"P123_omp_num_threads", &o.P123_omp_num_threads},
        {"PSI_start", &o.PSI_start}, {"PSI_end", &o.PSI_end}, {"channel_idx", &o.channel_idx}};
    const std::map<std::string, double *> dbls = {
        {"chebyshev_t", &o.chebyshev_t}, {"chebyshev_s", &o.chebyshev_s}, {"solve_tol", &o.solve_tol}};
    const std::map<std::string, bool *> bools = {
        {"parallel_run", &o.parallel_run}, {"P123_recovery", &o.P123_recovery},
        {"tensor_force", &o.tensor_force}, {"isospin_breaking_1S0", &o.isospin_breaking_1S0},
        {"midpoint_approx", &o.midpoint_approx}, {"calculate_and_store_P123", &o.calculate_and_store_P123},
        {"include_breakup_channels", &o.include_breakup_channels}, {"solve_faddeev", &o.solve_faddeev},
        {"solve_dense", &o.solve_dense}, {"production_run", &o.production_run},
        {"parameter_walk", &o.parameter_walk}};
    const std::map<std::string, std::string *> strs = {
        {"potential_model", &o.potential_model}, {"subfolder", &o.subfolder},
        {"p_grid_type", &o.p_grid_type}, {"p_grid_filename", &o.p_grid_filename},
        {"q_grid_type", &o.q_grid_type}, {"q_grid_filename", &o.q_grid_filename},
        {"parameter_file", &o.parameter_file}, {"energy_input_file", &o.energy_input_file},
        {"output_folder", &o.output_folder}, {"P123_folder", &o.P123_folder}, {"solve_method", &o.solve_method}};
    try {
        if (auto it = ints.find(option); it != ints.end()) {
            *it->second = std::stoi(input);
            return true;
        }
        if (auto it = dbls.find(option); it != dbls.end()) {
            *it->second = std::stod(input);
            return true;
        }
    } catch (const std::exception &) {
        die("Invalid value for input parameter " + option + "!");
    }
    if (auto it = bools.find(option); it != bools.end()) {
        *it->second = parse_bool(option, input);
        return true;
    }
    if (auto it = strs.find(option); it != strs.end()) {
        *it->second = input;
        return true;
    }
    return false;
}

void unrecognised_option(const std::string &s) {
    std::cout << "Unrecognised run option entered: " << s << "\nUse -h or --help for run options.\n\n"
              << "Skipping unrecognised run option.\n"
              << std::endl;
}

// read_input_list_and_set_parameters (set_run_parameters.cpp:253-285)
void read_input_list(Options &o, const std::string &filename) {
    std::ifstream in(filename);
    if (!in) die("Unable to open file " + filename);
    std::string line;
    while (std::getline(in, line)) {
        const size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        const std::string option = line.substr(0, eq);
        std::string input = line.substr(eq + 1);
        if (input.find_first_not_of(' ') == std::string::npos) continue;
        std::string stripped;
        for (char c : input)
            if (!std::isspace((unsigned char)c)) stripped += c;
        if (!set_option(o, option, stripped)) unrecognised_option(line);
    }
    std::cout << "Using input file " << filename << " for run options.\n\n"
              << "Any unset options will be filled with default parameters.\n"
              << "All other command line input will be ignored.\n"
              << std::endl;
}

void show_usage() {
    std::cout << "tictac_run [<inputfile>.txt] [option=value ...]   (-h or --help: this text)\n\n"
                 "Options (the reference's; the last value read wins):\n"
                 "  two_J_3N_max Np_WP Nq_WP J_2N_max Nphi Nx Np_per_WP Nq_per_WP chebyshev_s chebyshev_t\n"
                 "  tensor_force isospin_breaking_1S0 midpoint_approx production_run include_breakup_channels\n"
                 "  solve_faddeev solve_dense parallel_run channel_idx potential_model (malfliet_tjon, synthetic)\n"
                 "  parameter_walk parameter_file PSI_start PSI_end energy_input_file output_folder\n"
                 "  calculate_and_store_P123 P123_folder P123_recovery P123_omp_num_threads (the HDF5 P123 files)\n"
                 "  accepted and ignored: subfolder; p_grid_type/q_grid_type must be chebyshev\n"
                 "Added: solve_method=pade|dense|gmres|mixed (default pade, as the reference), solve_tol=1e-12\n"
              << std::endl;
}

// set_run_parameters (set_run_parameters.cpp:617-687)
Options parse_args(int argc, char **argv) {
    Options o;
    if (argc == 1) std::cout << "No run options entered.\nUse -h or --help for run options.\n\nRunning default setup.\n\n";
    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        if (arg == "-h" || arg == "--help") {
            show_usage();
            if (argc == 2) {
                std::cout << "Only -h or --help was requested. \nProgram will not run. Exiting ...\n" << std::endl;
                std::exit(0);
            }
        } else if (arg.size() >= 4 && arg.substr(arg.size() - 4) == ".txt") {
            read_input_list(o, arg);
        } else if (arg.find('=') != std::string::npos) {
            const size_t eq = arg.find('=');
            const std::string input = arg.substr(eq + 1);
            if (input.empty()) continue;
            if (!set_option(o, arg.substr(0, eq), input)) unrecognised_option(arg);
        } else {
            unrecognised_option(arg);
        }
    }
    if (o.parallel_run && o.channel_idx == -1)
        die("Parallel run specified in input but no channel index (i.e. channel_idx=<int>) given!");
    if (o.two_J_3N_max % 2 == 0 || o.two_J_3N_max <= 0) die("Cannot have even two_J_3N_max!");
    if (o.p_grid_type != "chebyshev" || o.q_grid_type != "chebyshev")
        die("only the chebyshev WP grid is built here (p_grid_type, q_grid_type)");
    return o;
}

// read_parameter_sample_list (disk_io_routines.cpp:43-78): one parameter set per line
std::vector<std::vector<double>> read_parameter_file(const std::string &filename) {
    std::ifstream in(filename);
    if (!in) die("Unable to open file " + filename);
    std::vector<std::vector<double>> sets;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::vector<double> row;
        std::string tok;
        while (ss >> tok) {
            try {
                row.push_back(std::stod(tok));
            } catch (const std::exception &) {
                die("unreadable value '" + tok + "' in " + filename);
            }
        }
        if (row.empty()) continue;
        if (!sets.empty() && row.size() != sets[0].size()) die("parameter sets of different lengths in " + filename);
        sets.push_back(row);
    }
    return sets;
}

// read_input_energies (disk_io_routines.cpp:80-110)
std::vector<double> read_energies(const std::string &path) {
    if (path.size() < 4 || path.substr(path.size() - 4) != ".txt") die("Input energy file must be .txt");
    std::ifstream in(path);
    if (!in) die("Unable to open file " + path);
    std::vector<double> e;
    double x;
    while (in >> x) e.push_back(x);
    return e;
}

// to_string_with_precision_and_sign (disk_io_routines.cpp:3-12)
std::string signed_sci(double v, int n = 16) {
    std::ostringstream out;
    out.precision(n);
    out << (v >= 0 ? "+" : "") << std::scientific << v;
    return out.str();
}

const char *kRule =
    "# ################################################################################################### \n";

struct ChannelData {
    const tt_pw_states *chn;
    const tt_channel_os *os;
    const tt_swp_states *swp;
    const tt_fwp_states *fwp;
    const tt_run_params *rp;
    std::string potential_model;
};

void write_header(std::ofstream &f, const char *kind, const ChannelData &c) {
    const int P = c.chn->P_3N[0];
    f << std::setprecision(8);
    f << "# " << kind << " Nd-scattering U-matrix elements in a wave-packet, Jj-scheme representation. \n";
    f << "# \n";
    f << "# The calculations below were done for: \n";
    f << "# JP:          " << c.chn->two_J_3N[0] << "/2" << (P == 1 ? "+" : "-") << "\n";
    f << "# Potential:   " << c.potential_model << "\n";
    f << "# Np:          " << c.swp->Np_WP << " \n";
    f << "# Nq:          " << c.swp->Nq_WP << " \n";
    f << "# Particles:   " << "nd-scattering" << "\n";
    f << "# Deuteron BE: " << c.swp->E_bound << " MeV \n";
    f << "# \n";
}

// store_U_matrix_elements_txt (disk_io_routines.cpp:759-913)
void store_U_txt(const cdouble *U, const ChannelData &c, const std::string &filename) {
    std::ofstream f(filename);
    if (!f) die("Unable to open file " + filename);
    const tt_channel_os &os = *c.os;
    const int nd = os.num_deuteron_states, nq = os.num_T_lab;
    write_header(f, "Elastic", c);
    f << "# Symbol definitions: \n";
    f << "# Uij:   U-matrix element in MeV with row-idx i and col-idx j (NOTE in wave-packet representation!)  \n";
    f << "# l:     Spectator nucleon orbital angular momentum  \n";
    f << "# j:     Spectator nucleon total angular momentum  \n";
    f << "# Tlab:  Laboratory scattering/kinetic energy  \n";
    f << "# Ecm:   Centre-of-mass scattering/kinetic energy  \n";
    f << "# q_idx: Index of on-shell q-momentum bin correponding to Tlab/Ecm  \n";
    f << "# \n";
    f << kRule;
    f << std::right << std::scientific << std::setprecision(16);
    f << "#";
    for (const char *h : {"Name", "row-idx", "col-idx", "l'", "2*j'", "l", "2*j"}) f << std::right << std::setw(10) << h;
    f << "\n";
    for (int dr = 0; dr < nd; ++dr) {
        const int ar = os.deuteron_idx[dr];
        for (int dc = 0; dc < nd; ++dc) {
            const int ac = os.deuteron_idx[dc];
            f << " " << std::right << std::setw(10) << ("U" + std::to_string(dr) + std::to_string(dc));
            f << std::setw(10) << dr << std::setw(10) << dc << std::setw(10) << c.chn->L_1N[ar] << std::setw(10)
              << c.chn->two_J_1N[ar] << std::setw(10) << c.chn->L_1N[ac] << std::setw(10) << c.chn->two_J_1N[ac]
              << "\n";
        }
    }
    f << kRule;
    const char *h2[] = {"Tlab [MeV]", "Ecm [MeV]", "q_idx",     "U00 [MeV]", "U01 [MeV]", "U02 [MeV]",
                        "U10 [MeV]",  "U11 [MeV]", "U12 [MeV]", "U20 [MeV]", "U21 [MeV]", "U22 [MeV]"};
    f << "#";
    for (int i = 0; i < 12; ++i) f << std::right << std::setw(i < 2 ? 24 : i == 2 ? 10 : 50) << h2[i];
    f << "\n";
    for (int j = 0; j < nq; ++j) {
        f << " " << std::right << std::setw(24) << os.T_lab[j] << std::setw(24) << os.E_com[j] << std::setw(10)
          << os.q_com_idx[j];
        for (int dr = 0; dr < nd; ++dr)
            for (int dc = 0; dc < nd; ++dc) {
                const cdouble u = U[(size_t)dr * nd * nq + (size_t)dc * nq + j];
                f << std::right << std::setw(50) << (signed_sci(u.real()) + signed_sci(u.imag()) + "j");
            }
        f << "\n";
    }
    f << kRule;
    f << std::endl;
}

// restructure_breakup_data + store_U_BU_matrix_elements_txt (disk_io_routines.cpp:341-706):
// one column group per distinct (breakup alpha', deuteron alpha) pair, one row per breakup cell
// of that pair, the rows of each energy after those of the previous one.  Bin edges of a
// neighbour cell outside the grid are written as NaN (the reference reads outside its arrays).
void store_U_BU_txt(const cdouble *U_BU, const ChannelData &c, const std::vector<double> &pe,
                    const std::string &filename) {
    std::ofstream f(filename);
    if (!f) die("Unable to open file " + filename);
    const tt_channel_os &os = *c.os;
    const tt_pw_states &w = *c.chn;
    const int nd = os.num_deuteron_states, nq = os.num_T_lab, Na = w.Nalpha;
    const int Np = c.swp->Np_WP, Nq = c.swp->Nq_WP;
    const double mu23 = 938.272 * 939.565 / (939.565 + 938.272);   // CPP/constants.h
    auto p_of = [&](double E) { return std::sqrt(2 * E * mu23); };
    auto q_edge = [&](int q) { return (q < 0 || q > Nq) ? NAN : c.fwp->q_WP[q]; };
    auto e_edge = [&](int a, int p) { return (p < 0 || p > Np) ? NAN : pe[(size_t)a * (Np + 1) + p]; };
    std::vector<std::string> headers = {"real", "Tlab [MeV]",        "real", "Ecm [MeV]",        "real",
                                        "qcm [MeV] (init.)", "real", "qlo [MeV] (init.)", "real",
                                        "qup [MeV] (init.)"};
    std::map<int, int> alpha_map, row_idx_map;
    std::vector<int> num_rows = {0};
    int alpha_counter = 0;
    for (int j = 0; j < nq; ++j) {
        std::map<int, int> row_num_map;
        for (int dr = 0; dr < nd; ++dr) {
            const int ar = os.deuteron_idx[dr];
            const std::string l = std::to_string(w.L_1N[ar]), jj = std::to_string(w.two_J_1N[ar]);
            for (int b = os.q_com_BU_idx[j]; b < os.q_com_BU_idx[j + 1]; ++b) {
                const int ac = os.alphapq_idx[3 * b];
                const std::string label = "<" + std::to_string(w.L_2N[ac]) + "," + std::to_string(w.S_2N[ac]) + "," +
                                          std::to_string(w.J_2N[ac]) + "," + std::to_string(w.L_1N[ac]) + "," +
                                          std::to_string(w.two_J_1N[ac]) + "|U|" + l + "," + jj + ">";
                const int key_alpha = ac * Na + ar, key_row = b * Na + ar;
                if (alpha_map.find(key_alpha) == alpha_map.end()) {
                    for (const char *h : {"comp", "", "real", "qlo [MeV]", "real", "qup [MeV]", "real", "plo [MeV]",
                                          "real", "pup [MeV]"})
                        headers.push_back(h);
                    headers[headers.size() - 9] = label;
                    alpha_map[key_alpha] = alpha_counter++;
                }
                if (row_num_map.find(key_alpha) == row_num_map.end()) {
                    row_idx_map[key_row] = num_rows[j];
                    row_num_map[key_alpha] = num_rows[j] + 1;
                } else {
                    row_idx_map[key_row] = row_num_map[key_alpha];
                    row_num_map[key_alpha] += 1;
                }
            }
        }
        int rows_q = 0;
        for (const auto &kv : row_num_map) rows_q = std::max(rows_q, kv.second);
        num_rows.push_back(rows_q);
    }
    const int n_rows = num_rows.back(), n_cols = (int)headers.size() / 2;
    std::vector<cdouble> data((size_t)n_rows * n_cols, cdouble(0, 0));
    for (int j = 0; j < nq; ++j) {
        const int qi = os.q_com_idx[j];
        for (int r = num_rows[j]; r < num_rows[j + 1]; ++r) {
            data[(size_t)r * n_cols + 0] = os.T_lab[j];
            data[(size_t)r * n_cols + 1] = os.E_com[j];
            data[(size_t)r * n_cols + 2] = os.q_com[j];
            data[(size_t)r * n_cols + 3] = c.fwp->q_WP[qi];
            data[(size_t)r * n_cols + 4] = c.fwp->q_WP[qi + 1];
        }
        for (int dr = 0; dr < nd; ++dr) {
            const int ar = os.deuteron_idx[dr];
            for (int b = os.q_com_BU_idx[j]; b < os.q_com_BU_idx[j + 1]; ++b) {
                const int ac = os.alphapq_idx[3 * b], qc = os.alphapq_idx[3 * b + 1], pc = os.alphapq_idx[3 * b + 2];
                const int col = 5 * alpha_map[ac * Na + ar] + 5, r = row_idx_map[b * Na + ar];
                cdouble *d = &data[(size_t)r * n_cols + col];
                d[0] = U_BU[(size_t)dr * os.num_BU_chns + b];
                d[1] = q_edge(qc);
                d[2] = q_edge(qc + 1);
                d[3] = p_of(e_edge(ac, pc));
                d[4] = p_of(e_edge(ac, pc + 1));
            }
        }
    }
    write_header(f, "Breakup", c);
    f << "# Symbol definitions: \n";
    f << "# <L,S,J,2*l,2*j|U|2*l,2*j>:   U-matrix element in MeV with for channel <L S J l j|U|L=0 S=1 J=1 l j> (NOTE in wave-packet representation!)  \n";
    f << "# L:                           Pair orbital angular momentum  \n";
    f << "# S:                           Pair spin  \n";
    f << "# J:                           Pair total angular momentum  \n";
    f << "# l:                           Spectator orbital angular momentum  \n";
    f << "# j:                           Spectator total angular momentum  \n";
    f << "# Tlab:                        Laboratory scattering/kinetic energy  \n";
    f << "# Ecm:                         Centre-of-mass scattering/kinetic energy  \n";
    f << "# qcm:                         Spectator centre-of-mass momentum bin midpoint \n";
    f << "# qlo:                         Spectator centre-of-mass momentum bin lower boundary \n";
    f << "# qup:                         Spectator centre-of-mass momentum bin upper boundary \n";
    f << "# plo:                         Pair centre-of-mass momentum bin lower boundary \n";
    f << "# pup:                         Pair centre-of-mass momentum bin upper boundary \n";
    f << "# init.:                       Pre-scattering \n";
    f << "# \n";
    f << kRule;
    f << std::right << std::scientific << std::setprecision(16);
    f << "#";
    for (int i = 0; i < n_cols; ++i) f << std::right << std::setw(headers[2 * i] == "real" ? 24 : 50) << headers[2 * i + 1];
    f << "\n";
    for (int r = 0; r < n_rows; ++r) {
        f << " ";
        for (int i = 0; i < n_cols; ++i) {
            const cdouble v = data[(size_t)r * n_cols + i];
            if (headers[2 * i] == "real")
                f << std::right << std::setw(24) << v.real();
            else
                f << std::right << std::setw(50) << (signed_sci(v.real()) + signed_sci(v.imag()) + "j");
        }
        f << "\n";
    }
    f << kRule;
    f << std::endl;
}

void make_dir(const std::string &path) {
    std::string cur;
    for (size_t i = 0; i <= path.size(); ++i) {
        if (i == path.size() || path[i] == '/') {
            if (!cur.empty() && mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
                die("cannot create output folder " + cur + ": " + std::strerror(errno));
        }
        if (i < path.size()) cur += path[i];
    }
}

}  // namespace

int main(int argc, char **argv) {
    const Options o = parse_args(argc, argv);
    if (!o.solve_faddeev && !o.calculate_and_store_P123)
        die("Both solve_faddeev=false and calculate_and_store_P123==false so code has nothing to do!");
    int method = TT_SOLVE_PADE;
    if (o.solve_method == "dense")
        method = TT_SOLVE_DENSE;
    else if (o.solve_method == "gmres")
        method = TT_SOLVE_GMRES;
    else if (o.solve_method == "mixed")
        method = TT_SOLVE_MIXED;
    else if (o.solve_method != "pade")
        die("Invalid value for input parameter solve_method! (pade, dense, gmres, mixed)");
    if (o.solve_dense) method = TT_SOLVE_DENSE;

    // potential and parameter sets (main.cpp:117-137)
    if (o.potential_model != "malfliet_tjon" && o.potential_model != "synthetic")
        die("Invalid potential model entered: " + o.potential_model + " (this build: malfliet_tjon, synthetic)");
    std::vector<std::vector<double>> sets;
    if (o.parameter_walk || o.potential_model == "synthetic") {
        if (o.parameter_file == "none" || o.parameter_file.empty())
            die("parameter_file required (parameter_walk=true or potential_model=synthetic)");
        sets = read_parameter_file(o.parameter_file);
        if (sets.empty()) die("no parameter sets in " + o.parameter_file);
    }
    int psi_min = 0, psi_max = 1;
    if (o.parameter_walk) {
        psi_min = o.PSI_start;
        psi_max = o.PSI_end;
        if (psi_min < 0 || psi_max > (int)sets.size() || psi_max <= psi_min)
            die("PSI-range given is larger than parameter-set input file.");
    }
    tt_potential *pot = nullptr;
    check(tt_potential_create(o.potential_model.c_str(), &pot));
    if (!sets.empty()) check(tt_potential_update_parameters(pot, sets[psi_min].data(), (int)sets[psi_min].size()));

    tt_run_params rp;
    tt_run_params_default(&rp);
    rp.two_J_3N_max = o.two_J_3N_max;
    rp.J_2N_max = o.J_2N_max;
    rp.Np_WP = o.Np_WP;
    rp.Nq_WP = o.Nq_WP;
    rp.Np_per_WP = o.Np_per_WP;
    rp.Nphi = o.Nphi;
    rp.Nx = o.Nx;
    rp.chebyshev_s = o.chebyshev_s;
    rp.chebyshev_t = o.chebyshev_t;
    rp.tensor_force = o.tensor_force;
    rp.isospin_breaking_1S0 = o.isospin_breaking_1S0;
    rp.midpoint_approx = o.midpoint_approx;

    std::printf("Constructing 3N partial-wave basis ... \n");
    tt_pw_states pw;
    check(tt_construct_pw_states(&rp, &pw));
    tt_fwp_states fwp;
    check(tt_make_fwp_statespace(&rp, &fwp));
    const std::vector<double> T_in = read_energies(o.energy_input_file);
    if (o.solve_faddeev) make_dir(o.output_folder);
    tt_ctx *ctx = nullptr;
    check(tt_ctx_create(0, &ctx));
    int nu = 0, nc = 0;
    tt_num_2N_states(&rp, &nu, &nc);
    const int Np = rp.Np_WP;

    for (int chn = 0; chn < pw.N_chn_3N; ++chn) {
        if (o.parallel_run && chn != o.channel_idx) continue;
        tt_pw_states view;
        check(tt_pw_channel(&pw, chn, &view));
        const int two_J = view.two_J_3N[0], P3 = view.P_3N[0];
        std::printf("Working on 3N-channel J_3N=%d/2, PAR=%d (channel %d of %d) with %d partial-wave states \n", two_J,
                    P3, chn + 1, pw.N_chn_3N, view.Nalpha);
        // fill_P123_arrays with its file cache (make_permutation_matrix.cpp:1297-1403)
        tt_csc P123;
        if (o.calculate_and_store_P123) make_dir(o.P123_folder);
        check_h5(tt_fill_P123_arrays_h5(ctx, &rp, &fwp, &view, o.production_run, o.P123_folder.c_str(),
                                        o.calculate_and_store_P123, o.P123_recovery, o.P123_omp_num_threads, &P123,
                                        nullptr));
        if (!o.solve_faddeev) {   // the P123 precalculation run: the file is all it leaves
            tt_csc_free(&P123);
            continue;
        }
        for (int psi = psi_min; psi < psi_max; ++psi) {
            if (o.parameter_walk) {
                std::printf("   - Setting model parameters for PSI %d ... \n", psi);
                check(tt_potential_update_parameters(pot, sets[psi].data(), (int)sets[psi].size()));
            }
            std::vector<double> Vu(std::max(nu * Np * Np, 1)), Vc(std::max(nc * 4 * Np * Np, 1));
            check(tt_calculate_potential_matrices(&rp, &fwp, &pw, pot, Vu.data(), Vc.data()));
            std::vector<double> eu(std::max(nu * (Np + 1), 1)), ec(std::max(nc * 2 * (Np + 1), 1));
            std::vector<double> Cu(Vu.size()), Cc(Vc.size());
            tt_swp_states swp;
            std::memset(&swp, 0, sizeof(swp));
            swp.e_SWP_unco = eu.data();
            swp.e_SWP_coup = ec.data();
            swp.C_SWP_unco = Cu.data();
            swp.C_SWP_coup = Cc.data();
            check(tt_make_swp_states(&rp, &fwp, &pw, Vu.data(), Vc.data(), &swp));
            std::printf("   - Using E_bound = %.5f MeV \n", swp.E_bound);
            tt_channel_os os;
            check(tt_find_on_shell(&swp, &view, T_in.data(), (int)T_in.size(), &os));
            if (o.include_breakup_channels) check(tt_find_breakup_bins(&rp, &swp, &view, &os));
            const int nd = os.num_deuteron_states, nE = os.num_T_lab;
            std::vector<cdouble> U((size_t)std::max(nd * nd * nE, 1));
            std::vector<cdouble> U_BU((size_t)std::max(nd * os.num_BU_chns, 1));
            tt_solve_opts so;
            tt_solve_opts_default(&so);
            so.method = method;
            so.tol = o.solve_tol;
            so.resolvent_on_device = 1;
            if (o.include_breakup_channels) so.U_BU = reinterpret_cast<double *>(U_BU.data());
            if (nd > 0 && nE > 0) {
                std::printf(" - Solving Faddeev equation ... \n");
                check(tt_solve_faddeev_equations(ctx, &rp, reinterpret_cast<double *>(U.data()), nullptr, &P123,
                                                 Vu.data(), Vc.data(), &swp, &os, &view, &so, nullptr));
            }
            const std::string id = "_Np_" + std::to_string(Np) + "_Nq_" + std::to_string(rp.Nq_WP) + "_JP_" +
                                   std::to_string(two_J) + "_" + std::to_string(P3) + "_Jmax_" +
                                   std::to_string(rp.J_2N_max);
            const ChannelData cd{&view, &os, &swp, &fwp, &rp, o.potential_model};
            store_U_txt(U.data(), cd, o.output_folder + "/U_PW_elements" + id + "_PSI_" + std::to_string(psi) + ".txt");
            if (o.include_breakup_channels) {
                std::vector<double> pe((size_t)view.Nalpha * (Np + 1));
                check(tt_channel_p_edges(&rp, &swp, &view, pe.data()));
                store_U_BU_txt(U_BU.data(), cd, pe,
                               o.output_folder + "/U_PW_breakup_elements" + id + "_PSI_" + std::to_string(psi) + ".txt");
            }
            tt_channel_os_free(&os);
        }
        tt_csc_free(&P123);
    }
    tt_ctx_destroy(ctx);
    tt_potential_destroy(pot);
    tt_fwp_states_free(&fwp);
    tt_pw_states_free(&pw);
    return 0;
}
