// topo.h -- host-side system input: run.conf, LAMMPS data files, the
// Kremer-Grest melt generator and the restart writer.  Plain C++ (no HIP).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "fmd/fmd.h"

namespace fmd {

// Parsed run.conf + topology; the union of SimParams (src/core/types.cuh:45-86)
// and TopologyData (src/io/config.hpp:8-31).
struct Topo {
    // run.conf scalars
    int natoms_conf = 0;
    int ntypes = 0;
    float rc = 0.0f, skin = 0.0f, dt = 0.0f;
    int64_t nsteps = 0;
    int ensemble = FMD_LANGEVIN;
    float T_start = 0, T_stop = 0, Tdamp = 0, P_start = 0, P_stop = 0, Pdamp = 0;
    int chain = 3;
    uint64_t seed = 42;
    fmd_conf_extra extra{};
    std::string data_file, coords_file;
    // topology / state (indexed by atom order in the input)
    float L = 0.0f;
    std::vector<float> pos4;     // x, y, z, type bits
    std::vector<float> vel3;
    std::vector<int32_t> image3;
    std::vector<int32_t> mol;
    std::vector<int32_t> bonds, btypes;  // 2 per bond, 1 per bond
    std::vector<int32_t> angles;         // 4 per angle (i, j, k, type)
    std::vector<float> lj;               // ntypes^2 * (eps, sig)
    std::vector<float> bparams;          // 4 per bond type
    std::vector<float> aparams;          // 2 per angle type
    std::vector<float> masses;           // per type
    std::vector<int32_t> table_idx;      // ntypes^2, -1 = no table
    std::vector<int32_t> table_n;        // points per table
    std::vector<float> table_range;      // rmin, rmax per table
    std::vector<float> table_data;       // r, e, f per point

    int natoms() const { return (int)(pos4.size() / 4); }
};

// One table of a tabulated potential file (src/io/table_parser.cpp:7-87).
struct TableRead {
    int n = 0;
    float rmin = 0.0f, rmax = 0.0f;
    std::vector<float> rows;  // r, e, f per point
};
int32_t parse_table_file(const std::string& path, const std::string& keyword, TableRead& out);

// Errors are reported as (status, message); the message is kept per thread.
void set_host_error(const std::string& msg);
const char* host_error();

int32_t parse_conf(const std::string& path, Topo& t);           // config.cpp:10-213
int32_t parse_lammps_data(const std::string& path, Topo& t);    // lammps_data.cpp:18-229
int32_t finalize_topo(Topo& t);                                 // checks + defaults
void build_exclusions(int n, const std::vector<int32_t>& bonds, std::vector<int32_t>& off,
                      std::vector<int32_t>& lst);               // lammps_data.cpp:231-253
int32_t kg_melt(int nchains, int chain_len, double rho, uint64_t seed, Topo& t);
int32_t write_lammps_data(const Topo& t, const std::string& path, const float* pos4,
                          const float* vel3, const int32_t* image3, float L, int64_t step);
std::string restart_filename(const std::string& base, int64_t step);  // lammps_data.cpp:255-260
std::string restart_final_filename(const std::string& base);          // lammps_data.cpp:262-267
void fill_desc(Topo& t, fmd_system_desc& d);

}  // namespace fmd
