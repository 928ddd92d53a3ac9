// NRG reader and writer (include/mbx/read_nrg.h): a line-oriented parser over the grammar of
// src/io_tools/read_nrg.cpp:43-298, then bblock::System construction; the writer of
// src/io_tools/write_nrg.cpp.
#include "mbx/read_nrg.h"

#include <algorithm>
#include <cctype>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

namespace tools {

namespace {

using bblock::CUException;

struct Reader {
    std::istream& is;
    size_t lineno = 0;

    // next line; false at the end of the stream
    bool next(std::string& line) {
        if (!std::getline(is, line)) return false;
        ++lineno;
        return true;
    }
};

std::string lower(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return s;
}

// first whitespace-separated word of a line, lower-cased ("" for a blank line)
std::string keyword(const std::string& line) {
    std::istringstream ss(line);
    std::string w;
    ss >> w;
    return lower(w);
}

[[noreturn]] void fail(const char* func, const Reader& r, const std::string& what, const std::string& line) {
    std::ostringstream os;
    os << what << " in line " << r.lineno << " of the NRG file." << std::endl << "Found: " << line << std::endl;
    throw CUException(func, __FILE__, __LINE__, os.str());
}

// read_nrg.cpp:175-294: MONOMER <id> / atom lines / ENDMON, repeated until ENDMOL
void read_monomers(Reader& r, NrgSystem& sys) {
    std::string line;
    std::vector<size_t> molecule;
    while (true) {
        if (!r.next(line)) fail(__func__, r, "Expected MONOMER keyword or ENDMOL", "end of file");
        std::istringstream ss(line);
        std::string w, id;
        ss >> w >> id;
        w = lower(w);
        if (w == "endmol" && !molecule.empty()) break;
        if (w != "monomer" || id.empty())
            fail(__func__, r, "Expected MONOMER keyword, along with the monomer id,", line);
        NrgMonomer m;
        m.id = lower(id);
        while (true) {
            if (!r.next(line)) fail(__func__, r, "Expected coordinates or ENDMON", "end of file");
            // the first line after MONOMER is always an atom line (read_nrg.cpp:229-246): a
            // monomer without atoms is an error
            if (keyword(line) == "endmon" && !m.atoms.empty()) break;
            if (keyword(line) == "endmon") fail(__func__, r, "Expected coordinates or ENDMON", line);
            std::istringstream as(line);
            std::string name, x, y, z;
            if (!(as >> name >> x)) fail(__func__, r, "Expected coordinates or ENDMON", line);
            as >> y >> z;
            double v[3];
            try {
                v[0] = std::stod(x);
                v[1] = std::stod(y);
                v[2] = std::stod(z);
            } catch (const std::exception&) {
                fail(__func__, r, "Expected numerical coordinates", line);
            }
            m.atoms.push_back(name);
            m.xyz.insert(m.xyz.end(), v, v + 3);
        }
        molecule.push_back(sys.monomers.size());
        sys.monomers.push_back(std::move(m));
    }
    sys.molecules.push_back(std::move(molecule));
}

// read_nrg.cpp:82-136: SYSTEM, then MOLECULE blocks until ENDSYS
void read_system(Reader& r, const std::string& first, NrgSystem& sys) {
    if (keyword(first) != "system") fail(__func__, r, "Expected SYSTEM keyword", first);
    std::string line;
    while (true) {
        if (!r.next(line)) fail(__func__, r, "Expected MOLECULE keyword or ENDSYS", "end of file");
        const std::string w = keyword(line);
        if (w == "endsys" && !sys.molecules.empty()) return;
        if (w != "molecule") fail(__func__, r, "Expected MOLECULE keyword", line);
        read_monomers(r, sys);
    }
}

}  // namespace

void ParseNrg(std::istream& is, std::vector<NrgSystem>& systems) {
    Reader r{is};
    std::string line;
    // read_nrg.cpp:57-78: systems follow one another until an empty line or the end of the file
    if (!r.next(line)) fail(__func__, r, "Expected SYSTEM keyword", "end of file");
    while (true) {
        NrgSystem sys;
        read_system(r, line, sys);
        systems.push_back(std::move(sys));
        if (!r.next(line) || line.empty()) return;
    }
}

void ReadNrg(const char* filename, std::vector<bblock::System>& systems) {
    if (!filename) throw CUException(__func__, __FILE__, __LINE__, "null NRG file name");
    std::ifstream ifs(filename);
    if (!ifs)
        throw CUException(__func__, __FILE__, __LINE__,
                          std::string("Could not open NRG file ") + filename + " for reading.");
    std::vector<NrgSystem> parsed;
    ParseNrg(ifs, parsed);
    for (const NrgSystem& p : parsed) {
        bblock::System sys;
        for (const NrgMonomer& m : p.monomers) sys.AddMonomer(m.xyz, m.atoms, m.id);
        for (const auto& mol : p.molecules) sys.AddMolecule(mol);
        sys.Initialize();
        systems.push_back(std::move(sys));
    }
}

void SaveNrg(std::ostream& os, std::vector<bblock::System> systems, int digits) {
    for (size_t i = 0; i < systems.size(); ++i) {
        bblock::System& sys = systems[i];
        const std::vector<double> xyz = sys.GetRealXyz();
        const std::vector<std::string> names = sys.GetRealAtomNames();
        os << "SYSTEM " << i << std::endl;
        for (size_t j = 0; j < sys.GetNumMol(); ++j) {
            os << "MOLECULE " << i << "." << j << std::endl;
            for (size_t m : sys.GetMolecule(j)) {
                os << "MONOMER " << sys.GetMonId(m) << std::endl;
                const size_t first = sys.GetFirstInd(m), nat = sys.GetMonNumAt(m);
                for (size_t a = first; a < first + nat; ++a) {
                    os << std::scientific << std::setprecision(digits) << std::setw(5) << std::left << names[a];
                    for (int k = 0; k < 3; ++k) os << std::setw(20) << std::right << xyz[3 * a + k];
                    os << std::endl;
                }
                os << "ENDMON " << std::endl;
            }
            os << "ENDMOL " << std::endl;
        }
        os << "ENDSYS " << std::endl;
    }
}

void WriteNrg(const char* filename, std::vector<bblock::System> systems, int digits) {
    if (!filename) throw std::runtime_error("could not open NRG file for writting");
    std::ofstream ofs(filename);
    if (!ofs) throw std::runtime_error("could not open NRG file for writting");
    SaveNrg(ofs, std::move(systems), digits);
}

}  // namespace tools
