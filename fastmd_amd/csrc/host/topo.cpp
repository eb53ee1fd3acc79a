// topo.cpp -- run.conf parser, LAMMPS data reader/writer, 1-2 exclusions and
// the Kremer-Grest melt generator.  Same keys, atom styles and defaults as the
// reference's src/io/config.cpp and src/io/lammps_data.cpp; every failure is a
// status code with a message instead of an exception or exit().
#include "topo.h"

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <unordered_map>

namespace fmd {

static thread_local std::string g_host_err;
void set_host_error(const std::string& msg) { g_host_err = msg; }
const char* host_error() { return g_host_err.c_str(); }

static int32_t fail(int32_t code, const std::string& msg) {
    set_host_error(msg);
    return code;
}

static float type_bits(int t) { float w; std::memcpy(&w, &t, 4); return w; }
static int type_of(float w) { int t; std::memcpy(&t, &w, 4); return t; }

static void copy_name(char* dst, const std::string& s) {
    std::strncpy(dst, s.c_str(), 255);
    dst[255] = '\0';
}

// ---------------------------------------------------------------------------
// Tabulated potential file (src/io/table_parser.cpp:7-87): blank lines and
// lines starting with '#' are skipped; the table starts at the line whose first
// word is the keyword; the next line is "N n" optionally followed by
// "R rlo rhi"; then n lines "index r energy force".  The r column must
// increase; without R the range is the first and last r.
// ---------------------------------------------------------------------------
int32_t parse_table_file(const std::string& path, const std::string& keyword, TableRead& out) {
    std::ifstream in(path);
    if (!in.is_open()) return fail(FMD_ERR_IO, "Cannot open table file: " + path);
    std::string line;
    bool found = false, has_r = false;
    int n = 0;
    float rlo = 0.0f, rhi = 0.0f;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        std::string first;
        iss >> first;
        if (first != keyword) continue;
        found = true;
        while (std::getline(in, line)) {
            if (line.empty() || line[0] == '#') continue;
            std::istringstream hs(line);
            std::string k;
            hs >> k;
            if (k == "N") {
                hs >> n;
                std::string rk;
                if (hs >> rk && rk == "R") {
                    hs >> rlo >> rhi;
                    has_r = true;
                }
            }
            break;
        }
        break;
    }
    if (!found) return fail(FMD_ERR_PARSE, "Table keyword not found: " + keyword);
    if (n < 2) return fail(FMD_ERR_PARSE, "table needs N >= 2 points: " + keyword);
    out.rows.clear();
    while ((int)(out.rows.size() / 3) < n && std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        int idx;
        float r, e, f;
        if (!(iss >> idx >> r >> e >> f)) return fail(FMD_ERR_PARSE, "bad table line in " + path);
        out.rows.push_back(r);
        out.rows.push_back(e);
        out.rows.push_back(f);
    }
    if ((int)(out.rows.size() / 3) != n) return fail(FMD_ERR_PARSE, "Table data line count mismatch in " + path);
    for (int i = 1; i < n; i++)
        if (!(out.rows[3 * i] > out.rows[3 * (i - 1)]))
            return fail(FMD_ERR_PARSE, "Table r values not monotonically increasing");
    if (!has_r) {
        rlo = out.rows[0];
        rhi = out.rows[3 * (n - 1)];
    }
    out.n = n;
    out.rmin = rlo;
    out.rmax = rhi;
    return FMD_OK;
}

// ---------------------------------------------------------------------------
// run.conf: one "key values..." per line, '#' in column 0 starts a comment,
// unknown keys are ignored (src/io/config.cpp:26-128).
// ---------------------------------------------------------------------------
int32_t parse_conf(const std::string& path, Topo& t) {
    std::ifstream in(path);
    if (!in.is_open()) return fail(FMD_ERR_IO, "Cannot open config: " + path);
    t.extra.restart_freq = -1;  // src/io/config.cpp:12
    struct LJE { int ti, tj; float eps, sig; };
    std::vector<LJE> lj;
    std::vector<std::array<float, 4>> btyp;
    std::vector<std::array<float, 2>> atyp;
    struct TabE { int ti, tj; std::string file, keyword; };
    std::vector<TabE> tabs;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        std::string key;
        iss >> key;
        if (key == "natoms") iss >> t.natoms_conf;
        else if (key == "ntypes") iss >> t.ntypes;
        else if (key == "rc") iss >> t.rc;
        else if (key == "skin") iss >> t.skin;
        else if (key == "dt") iss >> t.dt;
        else if (key == "temperature")
            return fail(FMD_ERR_PARSE, "'temperature' is deprecated. Use 'nvt_langevin T_start T_stop Tdamp [seed]' instead.");
        else if (key == "gamma")
            return fail(FMD_ERR_PARSE, "'gamma' is deprecated. Use 'nvt_langevin T_start T_stop Tdamp [seed]' instead.");
        else if (key == "seed")
            return fail(FMD_ERR_PARSE, "'seed' as a standalone key is deprecated. Specify seed in 'nvt_langevin T_start T_stop Tdamp [seed]'.");
        else if (key == "nsteps") { long long v = 0; iss >> v; t.nsteps = v; }
        else if (key == "dump_freq") iss >> t.extra.dump_freq;
        else if (key == "thermo") {
            std::string f;
            iss >> t.extra.thermo_on >> t.extra.thermo_freq >> f;
            copy_name(t.extra.thermo_file, f);
        } else if (key == "stress") {
            std::string f;
            iss >> t.extra.stress_on >> t.extra.stress_freq >> f;
            copy_name(t.extra.stress_file, f);
        } else if (key == "rg") {
            std::string f;
            iss >> t.extra.rg_on >> t.extra.rg_freq >> f;
            copy_name(t.extra.rg_file, f);
        } else if (key == "restart") {
            std::string f;
            iss >> t.extra.restart_freq >> f;
            copy_name(t.extra.restart_file, f);
        } else if (key == "nvt_langevin") {
            t.ensemble = FMD_LANGEVIN;
            iss >> t.T_start >> t.T_stop >> t.Tdamp;
            unsigned long long s;
            if (iss >> s) t.seed = s; else t.seed = 42;
        } else if (key == "nvt_nh") {
            t.ensemble = FMD_NVT_NH;
            iss >> t.T_start >> t.T_stop >> t.Tdamp;
            if (!(iss >> t.chain)) t.chain = 3;
        } else if (key == "npt_nh") {
            t.ensemble = FMD_NPT_NH;
            iss >> t.T_start >> t.T_stop >> t.Tdamp >> t.P_start >> t.P_stop >> t.Pdamp;
            if (!(iss >> t.chain)) t.chain = 3;
        } else if (key == "coords_file") iss >> t.coords_file;
        else if (key == "lammps_data_file") iss >> t.data_file;
        else if (key == "lj") {
            LJE e{0, 0, 0, 0};
            iss >> e.ti >> e.tj >> e.eps >> e.sig;
            lj.push_back(e);
        } else if (key == "table") {
            TabE e{0, 0, "", ""};
            iss >> e.ti >> e.tj >> e.file >> e.keyword;
            tabs.push_back(e);
        } else if (key == "bond_type") {
            int ty = 0; float k = 0, R0 = 0, eps = 0, sig = 0;
            iss >> ty >> k >> R0 >> eps >> sig;
            if (ty < 0 || ty > 1023) return fail(FMD_ERR_PARSE, "bond_type index out of range");
            if ((int)btyp.size() <= ty) btyp.resize(ty + 1, {0, 0, 0, 0});
            btyp[ty] = {k, R0, eps, sig};
        } else if (key == "angle_type") {
            int ty = 0; float k = 0, th = 0;
            iss >> ty >> k >> th;
            if (ty < 0 || ty > 1023) return fail(FMD_ERR_PARSE, "angle_type index out of range");
            if ((int)atyp.size() <= ty) atyp.resize(ty + 1, {0, 0});
            // degrees to radians in FP32 (src/io/config.cpp:176)
            atyp[ty] = {k, th * 3.14159265358979323846f / 180.0f};
        } else if (key == "forces_dump_file") {
            std::string f;
            iss >> f;
            copy_name(t.extra.forces_dump_file, f);
        }
    }
    if (t.ntypes < 0 || t.ntypes > 16) return fail(FMD_ERR_PARSE, "ntypes must be in [0, 16]");
    t.lj.assign(2 * (size_t)t.ntypes * t.ntypes, 0.0f);
    for (auto& e : lj) {
        if (e.ti < 0 || e.tj < 0 || e.ti >= t.ntypes || e.tj >= t.ntypes)
            return fail(FMD_ERR_PARSE, "lj type index out of range");
        t.lj[2 * (e.ti * t.ntypes + e.tj)] = e.eps; t.lj[2 * (e.ti * t.ntypes + e.tj) + 1] = e.sig;
        t.lj[2 * (e.tj * t.ntypes + e.ti)] = e.eps; t.lj[2 * (e.tj * t.ntypes + e.ti) + 1] = e.sig;
    }
    // tables per type pair; one table per (file, keyword) (src/io/config.cpp:136-162)
    t.table_idx.assign((size_t)t.ntypes * t.ntypes, -1);
    t.table_n.clear(); t.table_range.clear(); t.table_data.clear();
    std::vector<std::pair<std::string, std::string>> seen;
    for (auto& e : tabs) {
        if (e.ti < 0 || e.tj < 0 || e.ti >= t.ntypes || e.tj >= t.ntypes)
            return fail(FMD_ERR_PARSE, "table type index out of range");
        if (t.lj[2 * (e.ti * t.ntypes + e.tj)] != 0.0f || t.lj[2 * (e.ti * t.ntypes + e.tj) + 1] != 0.0f)
            return fail(FMD_ERR_PARSE, "Both lj and table defined for pair (" + std::to_string(e.ti) + "," +
                                           std::to_string(e.tj) + ")");
        int idx = -1;
        for (size_t k = 0; k < seen.size(); k++)
            if (seen[k].first == e.file && seen[k].second == e.keyword) idx = (int)k;
        if (idx < 0) {
            TableRead tr;
            if (int32_t rc = parse_table_file(e.file, e.keyword, tr)) return rc;
            idx = (int)seen.size();
            seen.emplace_back(e.file, e.keyword);
            t.table_n.push_back(tr.n);
            t.table_range.push_back(tr.rmin);
            t.table_range.push_back(tr.rmax);
            t.table_data.insert(t.table_data.end(), tr.rows.begin(), tr.rows.end());
        }
        t.table_idx[e.ti * t.ntypes + e.tj] = idx;
        t.table_idx[e.tj * t.ntypes + e.ti] = idx;
    }
    t.bparams.clear();
    for (auto& b : btyp) t.bparams.insert(t.bparams.end(), b.begin(), b.end());
    t.aparams.clear();
    for (auto& a : atyp) t.aparams.insert(t.aparams.end(), a.begin(), a.end());
    if (!t.coords_file.empty()) {  // xyz file, type 0 (src/io/config.cpp:179-192)
        std::ifstream cf(t.coords_file);
        if (!cf.is_open()) return fail(FMD_ERR_IO, "Cannot open coords_file: " + t.coords_file);
        int n = 0;
        cf >> n;
        std::string c;
        std::getline(cf, c);
        std::getline(cf, c);
        if (n < 0) return fail(FMD_ERR_PARSE, "bad atom count in coords_file");
        t.pos4.assign(4 * (size_t)n, 0.0f);
        t.vel3.assign(3 * (size_t)n, 0.0f);
        for (int i = 0; i < n; i++) {
            std::string el;
            float x, y, z;
            if (!(cf >> el >> x >> y >> z)) return fail(FMD_ERR_PARSE, "truncated coords_file");
            t.pos4[4 * i] = x; t.pos4[4 * i + 1] = y; t.pos4[4 * i + 2] = z; t.pos4[4 * i + 3] = type_bits(0);
        }
    }
    // ensemble validation (src/io/config.cpp:194-210)
    if (t.ensemble == FMD_LANGEVIN && !(t.Tdamp > 0.0f))
        return fail(FMD_ERR_PARSE, "Tdamp must be > 0 for nvt_langevin");
    if ((t.ensemble == FMD_NVT_NH || t.ensemble == FMD_NPT_NH)) {
        if (!(t.Tdamp > 0.0f)) return fail(FMD_ERR_PARSE, "Tdamp must be > 0 for nvt_nh/npt_nh");
        if (t.chain < 1) return fail(FMD_ERR_PARSE, "nh_chain_length must be >= 1");
    }
    if (t.ensemble == FMD_NPT_NH && !(t.Pdamp > 0.0f))
        return fail(FMD_ERR_PARSE, "Pdamp must be > 0 for npt_nh");
    if (!t.data_file.empty()) {
        // a relative data path is taken relative to the working directory, as in
        // the reference binary
        int32_t rc = parse_lammps_data(t.data_file, t);
        if (rc) return rc;
    }
    return finalize_topo(t);
}

// ---------------------------------------------------------------------------
// LAMMPS data file (src/io/lammps_data.cpp:18-229).
// ---------------------------------------------------------------------------
static bool section_header(const std::string& s) {
    return !s.empty() && std::isupper((unsigned char)s[0]);
}

int32_t parse_lammps_data(const std::string& path, Topo& t) {
    std::ifstream file(path);
    if (!file.is_open()) return fail(FMD_ERR_IO, "Cannot open LAMMPS data file: " + path);
    float Lx = 0, Ly = 0, Lz = 0;
    int nbox = 0;
    std::string line;
    while (nbox < 3 && std::getline(file, line)) {
        size_t s = line.find_first_not_of(" \t\r\n");
        if (s == std::string::npos || line[s] == '#') continue;
        float lo, hi;
        char lab[4];
        if (std::sscanf(line.c_str() + s, "%f %f %3s", &lo, &hi, lab) == 3) {
            std::string l(lab);
            if (l == "xlo") { Lx = hi - lo; nbox++; }
            else if (l == "ylo") { Ly = hi - lo; nbox++; }
            else if (l == "zlo") { Lz = hi - lo; nbox++; }
        }
    }
    if (nbox != 3) return fail(FMD_ERR_PARSE, "LAMMPS data file missing box dimensions");
    if (std::fabs(Lx - Ly) > 1e-6f || std::fabs(Lx - Lz) > 1e-6f)
        return fail(FMD_ERR_PARSE, "Non-cubic box not supported");
    t.L = Lx;
    file.clear();
    file.seekg(0);

    enum Sec { NONE, ATOMS, VELS, BONDS, ANGLES, MASSES } cur = NONE;
    // No "Atoms # style" hint: the style follows from the column count (5/8
    // atomic, 6/9 molecular, 7/10 full), which reads tests/fixtures/mini.data
    // the way tests/test_lammps_data.cu:10-21 expects.
    std::string style = "auto";
    std::unordered_map<long long, int> id2idx;
    std::vector<float> mass_in;
    t.pos4.clear(); t.vel3.clear(); t.image3.clear(); t.mol.clear();
    t.bonds.clear(); t.btypes.clear(); t.angles.clear();
    try {
        while (std::getline(file, line)) {
            size_t s = line.find_first_not_of(" \t\r\n");
            if (s == std::string::npos) continue;
            std::string tr = line.substr(s);
            if (tr[0] == '#') continue;
            if (section_header(tr)) {
                cur = NONE;
                if (tr.rfind("Atoms", 0) == 0) {
                    cur = ATOMS;
                    size_t h = tr.find('#');
                    if (h != std::string::npos) {
                        std::istringstream ss(tr.substr(h + 1));
                        ss >> style;
                    }
                } else if (tr.rfind("Velocities", 0) == 0) cur = VELS;
                else if (tr.rfind("Bonds", 0) == 0) cur = BONDS;
                else if (tr.rfind("Angles", 0) == 0) cur = ANGLES;
                else if (tr.rfind("Masses", 0) == 0) cur = MASSES;
                continue;
            }
            if (cur == NONE) continue;
            std::istringstream iss(tr);
            if (cur == ATOMS) {
                std::vector<std::string> tok;
                std::string x;
                while (iss >> x) tok.push_back(x);
                long long id; int mol, ty; float px, py, pz; int ix = 0, iy = 0, iz = 0;
                std::string st = style;
                if (st == "auto") {
                    const size_t nt = tok.size();
                    st = (nt == 5 || nt == 8) ? "atomic" : ((nt == 7 || nt == 10) ? "full" : "molecular");
                }
                if (st == "atomic") {
                    if (tok.size() < 5) continue;
                    id = std::stoll(tok[0]); mol = 1; ty = std::stoi(tok[1]);
                    px = std::stof(tok[2]); py = std::stof(tok[3]); pz = std::stof(tok[4]);
                    if (tok.size() >= 8) { ix = std::stoi(tok[5]); iy = std::stoi(tok[6]); iz = std::stoi(tok[7]); }
                } else if (st == "molecular") {
                    if (tok.size() < 6) continue;
                    id = std::stoll(tok[0]); mol = std::stoi(tok[1]); ty = std::stoi(tok[2]);
                    px = std::stof(tok[3]); py = std::stof(tok[4]); pz = std::stof(tok[5]);
                    if (tok.size() >= 9) { ix = std::stoi(tok[6]); iy = std::stoi(tok[7]); iz = std::stoi(tok[8]); }
                } else if (st == "full" || st == "charge") {
                    if (tok.size() < 7) continue;
                    id = std::stoll(tok[0]); mol = std::stoi(tok[1]); ty = std::stoi(tok[2]);
                    px = std::stof(tok[4]); py = std::stof(tok[5]); pz = std::stof(tok[6]);
                    if (tok.size() >= 10) { ix = std::stoi(tok[7]); iy = std::stoi(tok[8]); iz = std::stoi(tok[9]); }
                } else {  // unknown style: 6 fields id mol type x y z [ix iy iz]
                    if (tok.size() < 6) continue;
                    id = std::stoll(tok[0]); mol = std::stoi(tok[1]); ty = std::stoi(tok[2]);
                    px = std::stof(tok[3]); py = std::stof(tok[4]); pz = std::stof(tok[5]);
                    if (tok.size() >= 9) { ix = std::stoi(tok[6]); iy = std::stoi(tok[7]); iz = std::stoi(tok[8]); }
                }
                if (ty < 1 || ty > 16) return fail(FMD_ERR_PARSE, "atom type out of range [1,16] in " + path);
                id2idx[id] = t.natoms();
                t.mol.push_back(mol);
                t.pos4.push_back(px); t.pos4.push_back(py); t.pos4.push_back(pz);
                t.pos4.push_back(type_bits(ty - 1));
                t.image3.push_back(ix); t.image3.push_back(iy); t.image3.push_back(iz);
            } else if (cur == VELS) {
                long long id; float vx, vy, vz;
                if (!(iss >> id >> vx >> vy >> vz)) continue;
                auto it = id2idx.find(id);
                if (it == id2idx.end()) continue;
                if (t.vel3.size() < 3 * (size_t)t.natoms()) t.vel3.resize(3 * (size_t)t.natoms(), 0.0f);
                t.vel3[3 * it->second] = vx; t.vel3[3 * it->second + 1] = vy; t.vel3[3 * it->second + 2] = vz;
            } else if (cur == MASSES) {
                int ty; float m;
                if (!(iss >> ty >> m)) continue;
                if (m <= 0.0f) return fail(FMD_ERR_PARSE, "mass must be > 0 for type " + std::to_string(ty));
                if (ty < 1 || ty > 16) return fail(FMD_ERR_PARSE, "mass type out of range");
                if ((int)mass_in.size() < ty) mass_in.resize(ty, 1.0f);
                mass_in[ty - 1] = m;
            } else if (cur == BONDS) {
                long long id, a, b; int ty;
                if (!(iss >> id >> ty >> a >> b)) continue;
                auto i1 = id2idx.find(a), i2 = id2idx.find(b);
                if (i1 == id2idx.end() || i2 == id2idx.end()) continue;
                t.bonds.push_back(i1->second); t.bonds.push_back(i2->second);
                t.btypes.push_back(ty - 1);
            } else if (cur == ANGLES) {
                long long id, a, b, c; int ty;
                if (!(iss >> id >> ty >> a >> b >> c)) continue;
                auto i1 = id2idx.find(a), i2 = id2idx.find(b), i3 = id2idx.find(c);
                if (i1 == id2idx.end() || i2 == id2idx.end() || i3 == id2idx.end()) continue;
                t.angles.push_back(i1->second); t.angles.push_back(i2->second);
                t.angles.push_back(i3->second); t.angles.push_back(ty - 1);
            }
        }
    } catch (const std::exception& ex) {
        return fail(FMD_ERR_PARSE, std::string("malformed number in ") + path + ": " + ex.what());
    }
    int nt = 1;
    for (int i = 0; i < t.natoms(); i++) nt = std::max(nt, type_of(t.pos4[4 * i + 3]) + 1);
    mass_in.resize(std::max<size_t>(mass_in.size(), nt), 1.0f);
    t.masses = mass_in;
    return FMD_OK;
}

int32_t finalize_topo(Topo& t) {
    int n = t.natoms();
    if (n == 0) return fail(FMD_ERR_PARSE, "no atoms (give lammps_data_file or coords_file)");
    if (t.natoms_conf > 0 && t.natoms_conf != n)
        return fail(FMD_ERR_PARSE, "natoms " + std::to_string(t.natoms_conf) +
                                       " in run.conf does not match the " + std::to_string(n) +
                                       " atoms read");
    if (t.ntypes <= 0) t.ntypes = 1;
    for (int i = 0; i < n; i++) {
        int ty = type_of(t.pos4[4 * i + 3]);
        if (ty < 0 || ty >= t.ntypes) return fail(FMD_ERR_PARSE, "atom type outside ntypes");
    }
    if ((int)t.lj.size() != 2 * t.ntypes * t.ntypes) t.lj.assign(2 * (size_t)t.ntypes * t.ntypes, 0.0f);
    if (t.masses.size() < (size_t)t.ntypes) t.masses.resize(t.ntypes, 1.0f);
    if (t.vel3.size() < 3 * (size_t)n) t.vel3.resize(3 * (size_t)n, 0.0f);
    if (t.image3.size() < 3 * (size_t)n) t.image3.resize(3 * (size_t)n, 0);
    if (t.mol.size() < (size_t)n) t.mol.resize(n, 1);
    size_t nbt = t.bparams.size() / 4, nat = t.aparams.size() / 2;
    for (size_t b = 0; b < t.btypes.size(); b++)
        if (t.btypes[b] < 0 || (size_t)t.btypes[b] >= nbt)
            return fail(FMD_ERR_PARSE, "bond type without bond_type parameters");
    for (size_t a = 0; a < t.angles.size() / 4; a++)
        if (t.angles[4 * a + 3] < 0 || (size_t)t.angles[4 * a + 3] >= nat)
            return fail(FMD_ERR_PARSE, "angle type without angle_type parameters");
    if (!(t.L > 0.0f)) return fail(FMD_ERR_PARSE, "box side must be > 0");
    if (!(t.rc > 0.0f)) return fail(FMD_ERR_PARSE, "rc must be > 0");
    if (t.skin < 0.0f) return fail(FMD_ERR_PARSE, "skin must be >= 0");
    if (t.extra.thermo_on && t.extra.thermo_freq <= 0)
        return fail(FMD_ERR_PARSE, "thermo frequency must be > 0");
    if (t.extra.stress_on && t.extra.stress_freq <= 0)
        return fail(FMD_ERR_PARSE, "stress frequency must be > 0");
    if (t.extra.rg_on && t.extra.rg_freq <= 0) return fail(FMD_ERR_PARSE, "rg frequency must be > 0");
    return FMD_OK;
}

// 1-2 exclusions, CSR, sorted unique (src/io/lammps_data.cpp:231-253).
void build_exclusions(int n, const std::vector<int32_t>& bonds, std::vector<int32_t>& off,
                      std::vector<int32_t>& lst) {
    std::vector<std::vector<int32_t>> per(n);
    for (size_t b = 0; b + 1 < bonds.size(); b += 2) {
        per[bonds[b]].push_back(bonds[b + 1]);
        per[bonds[b + 1]].push_back(bonds[b]);
    }
    off.assign(n + 1, 0);
    lst.clear();
    for (int i = 0; i < n; i++) {
        auto& l = per[i];
        std::sort(l.begin(), l.end());
        l.erase(std::unique(l.begin(), l.end()), l.end());
        off[i] = (int32_t)lst.size();
        lst.insert(lst.end(), l.begin(), l.end());
    }
    off[n] = (int32_t)lst.size();
}

// ---------------------------------------------------------------------------
// Kremer-Grest melt generator.  Chains are random walks with bond length 0.97
// from a uniform random start.  Each new bead is drawn up to 64 times and kept
// at the first draw that is at least 1.02 from the bead two back and at least
// 0.75 from every other bead placed so far (checked on a host cell grid); when
// no draw qualifies the last one is kept.  Positions are wrapped into [0, L)
// with image counts.  Uniform variates are the top 53 bits of mt19937_64.
// ---------------------------------------------------------------------------
namespace {
struct HostGrid {
    int n;
    double cs, L;
    std::vector<std::vector<int>> cells;
    std::vector<double> xyz;  // wrapped coordinates of placed beads
    HostGrid(double L_, double cell) : L(L_) {
        n = std::max(1, (int)(L_ / cell));
        cs = L_ / n;
        cells.resize((size_t)n * n * n);
    }
    int cidx(double x) const {
        int c = (int)std::floor(x / cs);
        return ((c % n) + n) % n;
    }
    static double wrap(double x, double L) { return x - L * std::floor(x / L); }
    // min distance^2 from p to placed beads other than `skip`, capped at cut^2
    double min_d2(const double* p, int skip, double cut) const {
        const int cx = cidx(p[0]), cy = cidx(p[1]), cz = cidx(p[2]);
        const int r = std::min(n, (int)std::ceil(cut / cs));
        double best = cut * cut;
        for (int dz = -r; dz <= r; dz++)
            for (int dy = -r; dy <= r; dy++)
                for (int dx = -r; dx <= r; dx++) {
                    if (n < 2 * r + 1 && (dx != 0 || dy != 0 || dz != 0) &&
                        (std::abs(dx) > n / 2 || std::abs(dy) > n / 2 || std::abs(dz) > n / 2))
                        continue;
                    const auto& c = cells[(((cz + dz) % n + n) % n * n + ((cy + dy) % n + n) % n) * n +
                                          ((cx + dx) % n + n) % n];
                    for (int j : c) {
                        if (j == skip) continue;
                        double d2 = 0;
                        for (int a = 0; a < 3; a++) {
                            double d = p[a] - xyz[3 * (size_t)j + a];
                            d -= L * std::rint(d / L);
                            d2 += d * d;
                        }
                        if (d2 < best) best = d2;
                    }
                }
        return best;
    }
    void add(const double* p) {
        int id = (int)(xyz.size() / 3);
        for (int a = 0; a < 3; a++) xyz.push_back(p[a]);
        cells[((size_t)cidx(p[2]) * n + cidx(p[1])) * n + cidx(p[0])].push_back(id);
    }
};
}  // namespace

int32_t kg_melt(int nchains, int chain_len, double rho, uint64_t seed, Topo& t) {
    if (nchains < 1 || chain_len < 1 || !(rho > 0.0))
        return fail(FMD_ERR_ARG, "kg_melt: need nchains >= 1, chain_len >= 1, rho > 0");
    if ((int64_t)nchains * chain_len > (1 << 26)) return fail(FMD_ERR_ARG, "kg_melt: too many beads");
    std::mt19937_64 rng(seed);
    auto uni = [&]() { return (double)(rng() >> 11) * (1.0 / 9007199254740992.0); };
    const int n = nchains * chain_len;
    const double L = std::cbrt((double)n / rho);
    t = Topo();
    t.natoms_conf = n;
    t.ntypes = 1;
    t.rc = 2.5f; t.skin = 0.3f; t.dt = 0.001f; t.nsteps = 1000;
    t.ensemble = FMD_LANGEVIN; t.T_start = 1.0f; t.T_stop = 1.0f; t.Tdamp = 1.0f; t.seed = 42;
    t.extra.restart_freq = -1;
    t.L = (float)L;
    t.lj = {1.0f, 1.0f};
    t.bparams = {30.0f, 1.5f, 1.0f, 1.0f};
    t.aparams = {5.0f, 180.0f * 3.14159265358979323846f / 180.0f};
    t.masses = {1.0f};
    t.pos4.resize(4 * (size_t)n);
    t.vel3.assign(3 * (size_t)n, 0.0f);
    t.image3.assign(3 * (size_t)n, 0);
    t.mol.resize(n);
    const float Lf = (float)L;
    const double dmin = 0.75, dback = 1.02;
    HostGrid grid(L, 1.0);
    std::vector<double> un(3 * (size_t)chain_len);  // unwrapped chain coordinates
    for (int c = 0; c < nchains; c++) {
        for (int b = 0; b < chain_len; b++) {
            double* p = &un[3 * b];
            const int prev = c * chain_len + b - 1;
            for (int tries = 0; tries < 64; tries++) {
                if (b == 0) {
                    for (int d = 0; d < 3; d++) p[d] = uni() * L;
                } else {
                    const double* q = &un[3 * (b - 1)];
                    double z = 2.0 * uni() - 1.0, ph = 2.0 * M_PI * uni();
                    double s = std::sqrt(std::max(0.0, 1.0 - z * z));
                    p[0] = q[0] + 0.97 * s * std::cos(ph);
                    p[1] = q[1] + 0.97 * s * std::sin(ph);
                    p[2] = q[2] + 0.97 * z;
                }
                if (b >= 2) {
                    const double* q2 = &un[3 * (b - 2)];
                    double dx = p[0] - q2[0], dy = p[1] - q2[1], dz = p[2] - q2[2];
                    if (dx * dx + dy * dy + dz * dz < dback * dback) continue;
                }
                double w[3] = {HostGrid::wrap(p[0], L), HostGrid::wrap(p[1], L), HostGrid::wrap(p[2], L)};
                if (grid.min_d2(w, b > 0 ? prev : -1, dmin) >= dmin * dmin) break;
            }
            double w[3] = {HostGrid::wrap(p[0], L), HostGrid::wrap(p[1], L), HostGrid::wrap(p[2], L)};
            grid.add(w);
            int i = c * chain_len + b;
            for (int d = 0; d < 3; d++) {
                double im = std::floor(p[d] / L);
                float x = (float)(p[d] - im * L);
                if (x >= Lf) { x -= Lf; im += 1.0; }
                if (x < 0.0f) x = 0.0f;
                t.pos4[4 * i + d] = x;
                t.image3[3 * i + d] = (int32_t)im;
            }
            t.pos4[4 * i + 3] = type_bits(0);
            t.mol[i] = c + 1;
        }
        for (int b = 0; b + 1 < chain_len; b++) {
            t.bonds.push_back(c * chain_len + b);
            t.bonds.push_back(c * chain_len + b + 1);
            t.btypes.push_back(0);
        }
        for (int b = 0; b + 2 < chain_len; b++) {
            t.angles.push_back(c * chain_len + b);
            t.angles.push_back(c * chain_len + b + 1);
            t.angles.push_back(c * chain_len + b + 2);
            t.angles.push_back(0);
        }
    }
    return FMD_OK;
}

std::string restart_filename(const std::string& base, int64_t step) {
    size_t dot = base.rfind('.');
    if (dot != std::string::npos)
        return base.substr(0, dot) + "_" + std::to_string(step) + base.substr(dot);
    return base + "_" + std::to_string(step);
}
std::string restart_final_filename(const std::string& base) {
    size_t dot = base.rfind('.');
    if (dot != std::string::npos) return base.substr(0, dot) + "_final" + base.substr(dot);
    return base + "_final";
}

// Restart writer, the layout of write_lammps_data (src/io/lammps_data.cpp:269-385)
// with the zlo/zhi line written with its value.
int32_t write_lammps_data(const Topo& t, const std::string& path, const float* pos4,
                          const float* vel3, const int32_t* image3, float L, int64_t step) {
    FILE* fp = std::fopen(path.c_str(), "w");
    if (!fp) return fail(FMD_ERR_IO, "cannot open restart file " + path);
    const int n = t.natoms();
    const size_t nb = t.btypes.size(), na = t.angles.size() / 4;
    std::fprintf(fp, "# Restart from step %ld\n", (long)step);
    std::fprintf(fp, "%d atoms\n", n);
    std::fprintf(fp, "%d atom types\n", t.ntypes);
    if (nb > 0) { std::fprintf(fp, "%zu bonds\n", nb); std::fprintf(fp, "%zu bond types\n", t.bparams.size() / 4); }
    if (na > 0) { std::fprintf(fp, "%zu angles\n", na); std::fprintf(fp, "%zu angle types\n", t.aparams.size() / 2); }
    std::fprintf(fp, "\n");
    std::fprintf(fp, "0.000000 %.6f xlo xhi\n", L);
    std::fprintf(fp, "0.000000 %.6f ylo yhi\n", L);
    std::fprintf(fp, "0.000000 %.6f zlo zhi\n", L);
    std::fprintf(fp, "\nMasses\n\n");
    for (int i = 0; i < t.ntypes; i++)
        std::fprintf(fp, "%d %.6g\n", i + 1, i < (int)t.masses.size() ? t.masses[i] : 1.0f);
    std::fprintf(fp, "\nAtoms # full\n\n");
    for (int i = 0; i < n; i++) {
        int mol = i < (int)t.mol.size() ? t.mol[i] : 1;
        std::fprintf(fp, "%d %d %d 0.0 %.6f %.6f %.6f %d %d %d\n", i + 1, mol,
                     type_of(pos4[4 * i + 3]) + 1, pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2],
                     image3 ? image3[3 * i] : 0, image3 ? image3[3 * i + 1] : 0,
                     image3 ? image3[3 * i + 2] : 0);
    }
    std::fprintf(fp, "\nVelocities\n\n");
    for (int i = 0; i < n; i++)
        std::fprintf(fp, "%d %.6f %.6f %.6f\n", i + 1, vel3 ? vel3[3 * i] : 0.0f,
                     vel3 ? vel3[3 * i + 1] : 0.0f, vel3 ? vel3[3 * i + 2] : 0.0f);
    if (nb > 0) {
        std::fprintf(fp, "\nBonds\n\n");
        for (size_t b = 0; b < nb; b++)
            std::fprintf(fp, "%zu %d %d %d\n", b + 1, t.btypes[b] + 1, t.bonds[2 * b] + 1,
                         t.bonds[2 * b + 1] + 1);
    }
    if (na > 0) {
        std::fprintf(fp, "\nAngles\n\n");
        for (size_t a = 0; a < na; a++)
            std::fprintf(fp, "%zu %d %d %d %d\n", a + 1, t.angles[4 * a + 3] + 1,
                         t.angles[4 * a] + 1, t.angles[4 * a + 1] + 1, t.angles[4 * a + 2] + 1);
    }
    std::fclose(fp);
    return FMD_OK;
}

void fill_desc(Topo& t, fmd_system_desc& d) {
    std::memset(&d, 0, sizeof(d));
    d.natoms = t.natoms();
    d.ntypes = t.ntypes;
    d.pos = t.pos4.data();
    d.vel = t.vel3.empty() ? nullptr : t.vel3.data();
    d.image = t.image3.empty() ? nullptr : t.image3.data();
    d.mol_id = t.mol.empty() ? nullptr : t.mol.data();
    d.nbonds = (int32_t)t.btypes.size();
    d.bonds = t.bonds.empty() ? nullptr : t.bonds.data();
    d.bond_types = t.btypes.empty() ? nullptr : t.btypes.data();
    d.nangles = (int32_t)(t.angles.size() / 4);
    d.angles = t.angles.empty() ? nullptr : t.angles.data();
    d.masses = t.masses.empty() ? nullptr : t.masses.data();
    d.lj = t.lj.data();
    d.nbond_types = (int32_t)(t.bparams.size() / 4);
    d.bond_params = t.bparams.empty() ? nullptr : t.bparams.data();
    d.nangle_types = (int32_t)(t.aparams.size() / 2);
    d.angle_params = t.aparams.empty() ? nullptr : t.aparams.data();
    d.L = t.L;
    d.rc = t.rc;
    d.skin = t.skin;
    d.dt = t.dt;
    d.ensemble = t.ensemble;
    d.T_start = t.T_start; d.T_stop = t.T_stop; d.Tdamp = t.Tdamp;
    d.P_start = t.P_start; d.P_stop = t.P_stop; d.Pdamp = t.Pdamp;
    d.chain_length = t.chain;
    d.seed = t.seed;
    d.nsteps = t.nsteps;
    d.thermo_freq = t.extra.thermo_on ? t.extra.thermo_freq : 0;
    d.stress_freq = t.extra.stress_on ? t.extra.stress_freq : 0;
    d.force_cap = 0.0f;
    d.ntables = (int32_t)t.table_n.size();
    d.table_idx = d.ntables ? t.table_idx.data() : nullptr;
    d.table_n = d.ntables ? t.table_n.data() : nullptr;
    d.table_range = d.ntables ? t.table_range.data() : nullptr;
    d.table_data = d.ntables ? t.table_data.data() : nullptr;
}

}  // namespace fmd
