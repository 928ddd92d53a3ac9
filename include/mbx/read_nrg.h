// tools::ReadNrg -- the reference's NRG reader (src/io_tools/read_nrg.h:72-110,
// src/io_tools/read_nrg.cpp:43-298) over this library's bblock::System.
//
// File format (keywords case-insensitive, anything after a keyword on its line ignored):
//
//   SYSTEM
//   MOLECULE
//   MONOMER <id>
//   <atom name> <x> <y> <z>        (one line per atom, Angstrom)
//   ENDMON
//   MONOMER <id> ... ENDMON         (more monomers of the same molecule)
//   ENDMOL
//   MOLECULE ... ENDMOL             (more molecules)
//   ENDSYS
//   SYSTEM ... ENDSYS               (more systems; an empty line or the end of the file ends the list)
//
// Monomer ids are lower-cased (the reference's std::transform(::tolower)); atom names are kept as
// written.  Coordinates are read with std::stod, as the reference does.
#pragma once

#include <istream>
#include <ostream>
#include <string>
#include <vector>

#include "mbx/system.h"

namespace tools {

struct NrgMonomer {
    std::string id;
    std::vector<std::string> atoms;
    std::vector<double> xyz;  // 3 per atom
};

struct NrgSystem {
    std::vector<NrgMonomer> monomers;             // file order
    std::vector<std::vector<size_t>> molecules;   // monomer indices (into `monomers`) per molecule
};

// The reader of read_nrg.cpp:43-298 without the System: appends every system of `is` to
// `systems`.  A malformed file throws bblock::CUException naming the line (the reference's
// messages: "Expected SYSTEM keyword in line N", "Expected MOLECULE keyword ...", "Expected
// MONOMER keyword ..., along with the monomer id", "Expected coordinates or ENDMON ...",
// "Expected numerical coordinates ...").  Unlike the reference, a line that is neither MOLECULE
// nor ENDSYS after an ENDMOL is an error instead of being skipped.
void ParseNrg(std::istream& is, std::vector<NrgSystem>& systems);

// read_nrg.h:72.  Every system of the file is appended to `systems`: its monomers (AddMonomer)
// and molecules (AddMolecule) in file order, initialized (Initialize), as read_nrg.cpp:280,
// :293 and :135 do.  A System holds no device memory until its first computation, so reading a
// file of many configurations allocates nothing on the GPU.  A file that cannot be opened throws
// CUException("Could not open NRG file <name> for reading.").
void ReadNrg(const char* filename, std::vector<bblock::System>& systems);

// write_nrg.h:65-75 / write_nrg.cpp: every system as SYSTEM i / MOLECULE i.j / MONOMER <id> /
// atom lines (name left-aligned in 5 columns, coordinates in 20 columns, scientific, `digits`
// significant digits after the point: 8 in the reference) / ENDMON / ENDMOL / ENDSYS.  Molecules
// come from AddMolecule (a system without molecules writes none, as in the reference);
// monomers, names and coordinates (GetRealXyz: wrapped under PBC) are taken in input order.  The
// systems must be initialized.  Taken by value, as the reference does.  WriteNrg throws
// std::runtime_error("could not open NRG file for writting") -- the reference's text -- when the
// file cannot be opened.
void SaveNrg(std::ostream& os, std::vector<bblock::System> systems, int digits = 8);
void WriteNrg(const char* filename, std::vector<bblock::System> systems, int digits = 8);

}  // namespace tools
