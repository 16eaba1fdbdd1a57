// ingest.hpp -- conformational ensembles read from multi-model mmCIF / PDB
// text: the atom records of the first model and the coordinates of the
// selected atoms in every model, plus the PDB record writer used to write
// superposed models back out.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/bioshell_gpu.h"

namespace bs {

// One atom of the first model (bioshell_pdb::PdbAtom, pdb_atom.rs:20-50):
// strings are UTF-8; alt_loc and i_code hold one character.
struct AtomRecord {
    int32_t serial = 0;
    int32_t res_seq = 0;
    double occupancy = 0.0;
    double temp_factor = 0.0;
    std::string name;  // as stored by the reader (four characters)
    std::string alt_loc;
    std::string res_name;
    std::string chain_id;
    std::string i_code;
    std::string element;
    bool has_element = false;
    bool hetero = false;
};

// impl Display for PdbAtom (pdb_atom.rs:256-278), with the chain written as
// `chain` and the coordinates as given; no line terminator.
std::string pdb_atom_line(const AtomRecord& a, const std::string& chain, double x, double y,
                          double z);

// write_pdb (structure.rs:828-850) of one model: atoms grouped by chain in
// sorted chain order, chains renamed A..Z, 0..9 when any id is longer than one
// character.  xyz holds atoms.size() x 3 coordinates.
void append_pdb_model(std::string& out, const std::vector<AtomRecord>& atoms, const double* xyz);

}  // namespace bs

struct bs_ensemble {
    int format = 0;  // BS_FORMAT_CIF or BS_FORMAT_PDB
    size_t n_models = 0;
    size_t n_structure_atoms = 0;
    size_t n_chains = 0;
    size_t n_residues = 0;
    std::vector<bs::AtomRecord> atoms;  // the selected atoms, file order
    std::vector<double> xyz;            // n_models x atoms.size() x 3
};
