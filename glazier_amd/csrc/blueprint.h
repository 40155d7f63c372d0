// blueprint.h - the Blueprint JSON description, read in C++ for the CLI.
//
// Follows glazier-core/src/blueprint.rs: the same keys and serde defaults
// (blueprint.rs:15-193), the same checks (from_json, blueprint.rs:385-447),
// the adhesion fold into the contact matrix (effective_contact,
// blueprint.rs:200-239), the type assignment by fractions in label order
// (initial_types, blueprint.rs:247-261), and the model built from it
// (model, blueprint.rs:265-329). One extension: `initial.shuffle_seed`, when
// present, permutes that type list by a counter hash of the seed, which is how
// an infection is "uniformly placed by seed".
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/glazier_hip.h"

namespace ghip {

// A minimal JSON value: enough for Blueprint files (objects, arrays, numbers,
// strings, booleans, null).
struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;  // insertion order

    const Json *get(const std::string &key) const;
    bool is_number() const { return kind == Number; }
};

// Parses `text`; on failure returns false and sets `error`.
bool parse_json(const std::string &text, Json *out, std::string *error);

struct SpeciesSpec {
    std::string name;
    double diffusion = 0.0, decay = 0.0, initial = 0.0;
    int substeps = 0;  // extension: 0 takes the stability count
};

struct TypeSpec {
    std::string name;
    double target_volume = 0.0, lambda_volume = 0.0;
    std::map<std::string, double> secretion, uptake, chemotaxis, presents;
    double target_surface = 0.0, lambda_surface = 0.0;
    double division_volume = 0.0, death_rate = 0.0;
    double target_length = 0.0, lambda_length = 0.0;
    bool connected = false;
    double max_activity = 0.0, lambda_activity = 0.0;
    double external[3] = {0.0, 0.0, 0.0};
    double lambda_nematic = 0.0;
};

struct Blueprint {
    std::string name;
    uint32_t width = 0, height = 0, depth = 1;
    std::vector<TypeSpec> types;
    std::vector<SpeciesSpec> fields;
    bool has_adhesion = false;
    std::vector<std::string> molecules;
    std::vector<double> binding;
    std::vector<double> contact;
    double temperature = 0.0;
    uint32_t neighbour_order = 0;
    uint64_t seed = 0;
    uint64_t steps = 0, dump_every = 0;
    bool has_nematic_field = false;
    std::string nematic_field;
    uint32_t side = 0, nx = 0, ny = 0, nz = 1;
    std::vector<std::pair<std::string, double>> fractions;  // sorted by name, as a BTreeMap
    bool has_labels = false;
    std::string labels;
    bool has_shuffle = false;
    uint64_t shuffle_seed = 0;
    double micron_per_site = 0.0, minute_per_step = 0.0;

    std::vector<double> effective_contact() const;
    std::vector<uint8_t> initial_types(size_t cells) const;
    ghip_model model() const;
};

// Blueprint::from_json: parse, apply defaults, check names and the model.
bool blueprint_from_json(const std::string &text, Blueprint *out, std::string *error);

// Model::validate (glazier-core/src/model.rs:160-232) on the flat model.
bool validate_model(const ghip_model &m, std::string *error);

// Lattice::tile_grid (glazier-core/src/lattice.rs:169-196); returns the cell count.
uint32_t tile_grid(const ghip_model &m, uint32_t side, uint32_t nx, uint32_t ny, uint32_t nz,
                   std::vector<uint32_t> *labels);

// The `--check` reading (glazier/src/main.rs:59-101), as pretty JSON.
// What the reader took from a description, as `glazier --check` prints it
// (glazier/src/main.rs:62-96); with_depth adds the lattice depth, which the
// Python binding's parse reports beside it (glazier-py/src/lib.rs:18-150).
std::string check_json(const Blueprint &bp, bool with_depth = false);

// .npy I/O (glazier-core/src/npy.rs).
bool npy_write_u32(const std::string &path, const uint32_t *data, uint32_t w, uint32_t h, uint32_t d,
                   std::string *error);
bool npy_write_f64(const std::string &path, const double *data, uint32_t w, uint32_t h, uint32_t d,
                   std::string *error);
bool npy_read_labels(const std::string &path, std::vector<uint32_t> *out, std::string *error);
bool npy_read_pairs(const std::string &path, std::vector<double> *out, std::string *error);

}  // namespace ghip
