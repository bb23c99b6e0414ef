// nnpce — command-line driver with AtomNN's nnp_compute call surface, on the MI355X engine.
//
//   nnpce <param file>
//
// Parameter file (tags case-insensitive, '#' starts a comment; nnpce.cpp:149-166):
//   DATA   <file listing structure files, one per line>   (repeatable)
//   NNPOT  <ann potential file>
//   FORMAT XYZ | POSCAR
//   UNITS  METAL | AU
//   FORCE  T | F          print forces
//   NORM   T | F          print energies per atom
//
// Output follows nnpce.cpp:284-315: "energy <index> <E_ref> <E_nnp>" (or "energy/atom ...")
// per structure, then with FORCE a "struc[<index>]" block of "force[<i>] = fx fy fz" lines
// (the reference labels every line of a structure with the structure's loop index i).
// All structures are evaluated in one nnp_compute_batch call.
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nnp_capi.h"
#include "../../include/nnp_pot.hpp"

namespace {

// chem/units.hpp:10-16
const double kBohrPerAng = 1.0 / 0.529177210903;
const double kHartreePerEv = 1.0 / 27.211386245988;

struct Struc {
    std::vector<std::string> names;
    std::vector<double> pos;  // [n][3]
    double cell[9] = {0};     // lattice vectors a, b, c
    bool periodic = false;
    double energy = 0;
};

std::vector<std::string> split(const std::string& s) {
    std::vector<std::string> out;
    std::istringstream in(s);
    std::string t;
    while (in >> t) out.push_back(t);
    return out;
}

std::string upper(std::string s) {
    for (auto& c : s) c = (char)std::toupper((unsigned char)c);
    return s;
}

bool boolean(const std::string& s) {  // str/string.cpp:177-187
    return s == "T" || s == "t" || s == "TRUE" || s == "true" || s == "True" || s == "1";
}

bool getline(FILE* f, std::string& line) {
    line.clear();
    int ch;
    while ((ch = std::fgetc(f)) != EOF) {
        if (ch == '\n') return true;
        line.push_back((char)ch);
    }
    return !line.empty();
}

// format/xyz_struc.cpp:38-149 (orthorhombic cell on the comment line: "E lx ly lz")
Struc read_xyz(const std::string& path, double s_len, double s_energy) {
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) throw std::runtime_error("ERROR in XYZ::read(const char*,const AtomType&,Structure&): Could not open file: " + path);
    Struc s;
    std::string line;
    getline(f, line);
    const int n = std::atoi(line.c_str());
    if (n <= 0) {
        std::fclose(f);
        throw std::runtime_error("ERROR in XYZ::read(const char*,const AtomType&,Structure&): found zero atoms.");
    }
    getline(f, line);
    const auto tok = split(line);
    double lv[3] = {0, 0, 0};
    if (tok.size() == 2) {
        s.energy = std::atof(tok[1].c_str());
    } else if (tok.size() == 5) {
        s.energy = std::atof(tok[1].c_str());
        for (int d = 0; d < 3; ++d) lv[d] = std::atof(tok[2 + d].c_str());
    }
    for (int d = 0; d < 3; ++d) s.cell[4 * d] = lv[d] * s_len;  // a along x, b along y, c along z
    s.periodic = (lv[0] * lv[0] + lv[1] * lv[1] + lv[2] * lv[2]) > 0;
    for (int i = 0; i < n; ++i) {
        if (!getline(f, line)) {
            std::fclose(f);
            throw std::runtime_error("XYZ::read: unexpected end of file in " + path);
        }
        const auto t = split(line);
        if (t.size() < 4) {
            std::fclose(f);
            throw std::runtime_error("XYZ::read: malformed atom line in " + path);
        }
        s.names.push_back(t[0]);
        for (int d = 0; d < 3; ++d) s.pos.push_back(std::atof(t[1 + d].c_str()) * s_len);
    }
    std::fclose(f);
    s.energy *= s_energy;
    return s;
}

// format/vasp_struc.cpp:51-242.  Lattice vector lines are rows in the file and columns of R.
// Direct coordinates are multiplied by R; Cartesian coordinates by the length unit only, as the
// reference reader does (vasp_struc.cpp:159-168).
Struc read_poscar(const std::string& path, double s_len) {
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) throw std::runtime_error(std::string("I/O Error: Could not open file: ") + path);
    Struc s;
    std::string line;
    getline(f, line);  // title
    getline(f, line);
    const double scale = std::atof(line.c_str());
    for (int c = 0; c < 3; ++c) {
        getline(f, line);
        const auto t = split(line);
        if (t.size() < 3) {
            std::fclose(f);
            throw std::runtime_error("I/O Exception: Could not read data.");
        }
        for (int r = 0; r < 3; ++r) s.cell[3 * c + r] = std::atof(t[r].c_str()) * s_len * scale;
    }
    s.periodic = true;
    std::string names_line, counts_line;
    getline(f, names_line);
    getline(f, counts_line);
    const auto names = split(names_line), counts = split(counts_line);
    if (names.size() != counts.size() || names.empty()) {
        std::fclose(f);
        throw std::runtime_error("Invalid number of species");
    }
    getline(f, line);
    const bool direct = !line.empty() && line[0] == 'D';
    int n = 0;
    for (size_t k = 0; k < names.size(); ++k) {
        const int c = std::atoi(counts[k].c_str());
        for (int q = 0; q < c; ++q) s.names.push_back(names[k]);
        n += c;
    }
    for (int i = 0; i < n; ++i) {
        if (!getline(f, line)) {
            std::fclose(f);
            throw std::runtime_error("I/O Exception: Could not read data.");
        }
        const auto t = split(line);
        double v[3] = {0, 0, 0};
        for (int d = 0; d < 3 && d < (int)t.size(); ++d) v[d] = std::atof(t[d].c_str());
        if (direct) {
            for (int r = 0; r < 3; ++r)
                s.pos.push_back(s.cell[r] * v[0] + s.cell[3 + r] * v[1] + s.cell[6 + r] * v[2]);
        } else {
            for (int r = 0; r < 3; ++r) s.pos.push_back(v[r] * s_len);
        }
    }
    std::fclose(f);
    return s;
}

// Eigen's default row-vector print: coefficients right-aligned to a common width, one space apart.
std::string eigen_row(const double* v) {
    std::string c[3];
    size_t w = 0;
    for (int d = 0; d < 3; ++d) {
        std::ostringstream o;
        o << v[d];
        c[d] = o.str();
        w = std::max(w, c[d].size());
    }
    std::string out;
    for (int d = 0; d < 3; ++d) {
        if (d) out += " ";
        out += std::string(w - c[d].size(), ' ') + c[d];
    }
    return out;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc != 2) throw std::invalid_argument("Invalid number of command-line arguments.");
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<std::string> data;
        std::string pot, format, units;
        bool force = false, norm = false;
        {
            FILE* f = std::fopen(argv[1], "r");
            if (!f) throw std::runtime_error("I/O Error: could not open parameter file.");
            std::string line;
            while (getline(f, line)) {
                const size_t h = line.find('#');
                if (h != std::string::npos) line = line.substr(0, h);
                const auto t = split(line);
                if (t.empty()) continue;
                const std::string tag = upper(t[0]);
                const std::string val = t.size() > 1 ? t[1] : "";
                if (tag == "DATA") data.push_back(val);
                else if (tag == "NNPOT") pot = val;
                else if (tag == "FORMAT") format = upper(val);
                else if (tag == "UNITS") units = upper(val);
                else if (tag == "FORCE") force = boolean(val);
                else if (tag == "NORM") norm = boolean(val);
            }
            std::fclose(f);
        }
        double s_len = 1.0, s_energy = 1.0;
        if (units == "AU") {
            s_len = kBohrPerAng;
            s_energy = kHartreePerEv;
        } else if (units != "METAL" && units != "LJ") {
            throw std::invalid_argument("Invalid unit system.");
        }
        if (format != "XYZ" && format != "POSCAR") throw std::invalid_argument("Invalid file format.");
        if (data.empty()) throw std::invalid_argument("No data");
        std::vector<std::string> files;
        for (const auto& d : data) {
            FILE* f = std::fopen(d.c_str(), "r");
            if (!f) throw std::runtime_error("I/O Error: Could not open data file: " + d);
            std::string line;
            while (getline(f, line)) {
                const auto t = split(line);
                if (!t.empty()) files.push_back(t[0]);
            }
            std::fclose(f);
        }
        std::vector<Struc> strucs;
        for (const auto& fn : files)
            strucs.push_back(format == "XYZ" ? read_xyz(fn, s_len, s_energy) : read_poscar(fn, s_len));

        atomnn::GpuPot gpu(pot);
        const size_t ns = strucs.size();
        std::vector<std::vector<int32_t>> types(ns);
        std::vector<std::vector<double>> eatom(ns), forces(ns);
        std::vector<nnp_structure> in(ns);
        std::vector<double*> ea_ptr(ns), f_ptr;
        for (size_t k = 0; k < ns; ++k) {
            const Struc& s = strucs[k];
            const int n = (int)s.names.size();
            for (const auto& nm : s.names) {
                const int t = gpu.index(nm);
                if (t < 0) throw std::runtime_error("No match for key in map");
                types[k].push_back(t);
            }
            eatom[k].resize(n);
            forces[k].resize(3 * (size_t)n);
            in[k] = nnp_structure{n, types[k].data(), s.pos.data(), s.periodic ? s.cell : nullptr};
            ea_ptr[k] = eatom[k].data();
            if (force) f_ptr.push_back(forces[k].data());
        }
        std::vector<double> energy;
        gpu.compute_batch(in, energy, ea_ptr, f_ptr);
        for (size_t k = 0; k < ns; ++k) {
            const double n = (double)strucs[k].names.size();
            if (!norm) std::printf("energy %zu %.6f %.6f\n", k, strucs[k].energy, energy[k]);
            else std::printf("energy/atom %zu %.6f %.6f\n", k, strucs[k].energy / n, energy[k] / n);
        }
        std::fflush(stdout);
        if (force) {
            for (size_t k = 0; k < ns; ++k) {
                std::cout << "struc[" << k << "]\n";
                for (size_t j = 0; j < strucs[k].names.size(); ++j)
                    std::cout << "force[" << k << "] = " << eigen_row(&forces[k][3 * j]) << "\n";
            }
        }
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cout << "time = " << sec << "\n";
    } catch (std::exception& e) {
        std::cout << "ERROR in nn_pot_compute::main(int,char**):\n" << e.what() << "\n";
        return 1;
    }
    return 0;
}
