// blueprint.cpp - Blueprint JSON reader, model construction, .npy I/O.
// See blueprint.h for the reference lines each part follows.
#include "blueprint.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

#include "ghip_common.h"

namespace ghip {

// ---------------------------------------------------------------- JSON

const Json *Json::get(const std::string &key) const {
    if (kind != Object) return nullptr;
    for (const auto &kv : obj)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

namespace {

struct Parser {
    const std::string &s;
    size_t i = 0;
    std::string err;

    explicit Parser(const std::string &text) : s(text) {}

    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\r' || s[i] == '\t')) ++i;
    }
    bool fail(const std::string &m) {
        if (err.empty()) {
            size_t line = 1, col = 1;
            for (size_t k = 0; k < i && k < s.size(); ++k) {
                if (s[k] == '\n') {
                    ++line;
                    col = 1;
                } else {
                    ++col;
                }
            }
            err = m + " at line " + std::to_string(line) + " column " + std::to_string(col);
        }
        return false;
    }
    bool value(Json *out) {
        ws();
        if (i >= s.size()) return fail("EOF while parsing a value");
        char c = s[i];
        if (c == '{') return object(out);
        if (c == '[') return array(out);
        if (c == '"') {
            out->kind = Json::String;
            return string(&out->str);
        }
        if (c == 't' && s.compare(i, 4, "true") == 0) {
            i += 4;
            out->kind = Json::Bool;
            out->b = true;
            return true;
        }
        if (c == 'f' && s.compare(i, 5, "false") == 0) {
            i += 5;
            out->kind = Json::Bool;
            out->b = false;
            return true;
        }
        if (c == 'n' && s.compare(i, 4, "null") == 0) {
            i += 4;
            out->kind = Json::Null;
            return true;
        }
        return number(out);
    }
    bool number(Json *out) {
        size_t start = i;
        if (i < s.size() && (s[i] == '-' || s[i] == '+')) ++i;
        while (i < s.size() && (std::isdigit((unsigned char)s[i]) || s[i] == '.' || s[i] == 'e' ||
                                s[i] == 'E' || s[i] == '-' || s[i] == '+'))
            ++i;
        if (i == start) return fail("expected value");
        std::string tok = s.substr(start, i - start);
        char *end = nullptr;
        out->num = std::strtod(tok.c_str(), &end);
        if (!end || *end != '\0') return fail("invalid number");
        out->kind = Json::Number;
        return true;
    }
    bool string(std::string *out) {
        ++i;  // opening quote
        out->clear();
        while (i < s.size() && s[i] != '"') {
            if (s[i] == '\\') {
                ++i;
                if (i >= s.size()) return fail("EOF in string");
                char e = s[i];
                if (e == 'n') out->push_back('\n');
                else if (e == 't') out->push_back('\t');
                else if (e == 'r') out->push_back('\r');
                else if (e == 'b') out->push_back('\b');
                else if (e == 'f') out->push_back('\f');
                else if (e == 'u') {
                    if (i + 4 >= s.size()) return fail("bad unicode escape");
                    unsigned cp = (unsigned)std::strtoul(s.substr(i + 1, 4).c_str(), nullptr, 16);
                    i += 4;
                    if (cp < 0x80) out->push_back((char)cp);
                    else if (cp < 0x800) {
                        out->push_back((char)(0xC0 | (cp >> 6)));
                        out->push_back((char)(0x80 | (cp & 0x3F)));
                    } else {
                        out->push_back((char)(0xE0 | (cp >> 12)));
                        out->push_back((char)(0x80 | ((cp >> 6) & 0x3F)));
                        out->push_back((char)(0x80 | (cp & 0x3F)));
                    }
                } else out->push_back(e);
                ++i;
            } else {
                out->push_back(s[i++]);
            }
        }
        if (i >= s.size()) return fail("EOF while parsing a string");
        ++i;
        return true;
    }
    bool array(Json *out) {
        out->kind = Json::Array;
        ++i;
        ws();
        if (i < s.size() && s[i] == ']') {
            ++i;
            return true;
        }
        for (;;) {
            Json v;
            if (!value(&v)) return false;
            out->arr.push_back(std::move(v));
            ws();
            if (i < s.size() && s[i] == ',') {
                ++i;
                continue;
            }
            if (i < s.size() && s[i] == ']') {
                ++i;
                return true;
            }
            return fail("expected `,` or `]`");
        }
    }
    bool object(Json *out) {
        out->kind = Json::Object;
        ++i;
        ws();
        if (i < s.size() && s[i] == '}') {
            ++i;
            return true;
        }
        for (;;) {
            ws();
            if (i >= s.size() || s[i] != '"') return fail("key must be a string");
            std::string key;
            if (!string(&key)) return false;
            ws();
            if (i >= s.size() || s[i] != ':') return fail("expected `:`");
            ++i;
            Json v;
            if (!value(&v)) return false;
            out->obj.emplace_back(std::move(key), std::move(v));
            ws();
            if (i < s.size() && s[i] == ',') {
                ++i;
                continue;
            }
            if (i < s.size() && s[i] == '}') {
                ++i;
                return true;
            }
            return fail("expected `,` or `}`");
        }
    }
};

}  // namespace

bool parse_json(const std::string &text, Json *out, std::string *error) {
    Parser p(text);
    if (!p.value(out)) {
        *error = p.err;
        return false;
    }
    p.ws();
    if (p.i != text.size()) {
        p.fail("trailing characters");
        *error = p.err;
        return false;
    }
    return true;
}

// ---------------------------------------------------------------- typed reads

namespace {

struct Reader {
    std::string err;
    bool ok() const { return err.empty(); }
    void missing(const std::string &field) {
        if (err.empty()) err = "missing field `" + field + "`";
    }
    void wrong(const std::string &field, const std::string &want) {
        if (err.empty()) err = "invalid type for `" + field + "`, expected " + want;
    }
    double f64(const Json &o, const char *key, bool required, double dflt) {
        const Json *v = o.get(key);
        if (!v) {
            if (required) missing(key);
            return dflt;
        }
        if (!v->is_number()) {
            wrong(key, "f64");
            return dflt;
        }
        return v->num;
    }
    uint64_t uint(const Json &o, const char *key, bool required, uint64_t dflt) {
        const Json *v = o.get(key);
        if (!v) {
            if (required) missing(key);
            return dflt;
        }
        if (!v->is_number() || v->num < 0 || v->num != std::floor(v->num)) {
            wrong(key, "an unsigned integer");
            return dflt;
        }
        return (uint64_t)v->num;
    }
    std::string str(const Json &o, const char *key, bool required) {
        const Json *v = o.get(key);
        if (!v) {
            if (required) missing(key);
            return "";
        }
        if (v->kind != Json::String) {
            wrong(key, "a string");
            return "";
        }
        return v->str;
    }
    bool boolean(const Json &o, const char *key, bool dflt) {
        const Json *v = o.get(key);
        if (!v) return dflt;
        if (v->kind != Json::Bool) {
            wrong(key, "a boolean");
            return dflt;
        }
        return v->b;
    }
    std::map<std::string, double> map(const Json &o, const char *key) {
        std::map<std::string, double> out;
        const Json *v = o.get(key);
        if (!v) return out;
        if (v->kind != Json::Object) {
            wrong(key, "a map");
            return out;
        }
        for (const auto &kv : v->obj) {
            if (!kv.second.is_number()) wrong(key, "numbers");
            out[kv.first] = kv.second.num;
        }
        return out;
    }
    std::vector<double> numbers(const Json &o, const char *key, bool required) {
        std::vector<double> out;
        const Json *v = o.get(key);
        if (!v) {
            if (required) missing(key);
            return out;
        }
        if (v->kind != Json::Array) {
            wrong(key, "a sequence");
            return out;
        }
        for (const auto &x : v->arr) {
            if (!x.is_number()) wrong(key, "numbers");
            out.push_back(x.num);
        }
        return out;
    }
};

}  // namespace

bool blueprint_from_json(const std::string &text, Blueprint *out, std::string *error) {
    Json root;
    if (!parse_json(text, &root, error)) return false;
    if (root.kind != Json::Object) {
        *error = "a description is a JSON object";
        return false;
    }
    Reader r;
    Blueprint bp;
    bp.name = r.str(root, "name", true);
    bp.width = (uint32_t)r.uint(root, "width", true, 0);
    bp.height = (uint32_t)r.uint(root, "height", true, 0);
    bp.depth = (uint32_t)r.uint(root, "depth", false, 1);
    const Json *types = root.get("types");
    if (!types) r.missing("types");
    else if (types->kind != Json::Array) r.wrong("types", "a sequence");
    else
        for (const auto &t : types->arr) {
            TypeSpec ts;
            ts.name = r.str(t, "name", true);
            ts.target_volume = r.f64(t, "target_volume", true, 0.0);
            ts.lambda_volume = r.f64(t, "lambda_volume", true, 0.0);
            ts.secretion = r.map(t, "secretion");
            ts.uptake = r.map(t, "uptake");
            ts.chemotaxis = r.map(t, "chemotaxis");
            ts.presents = r.map(t, "presents");
            ts.target_surface = r.f64(t, "target_surface", false, 0.0);
            ts.lambda_surface = r.f64(t, "lambda_surface", false, 0.0);
            ts.division_volume = r.f64(t, "division_volume", false, 0.0);
            ts.death_rate = r.f64(t, "death_rate", false, 0.0);
            ts.target_length = r.f64(t, "target_length", false, 0.0);
            ts.lambda_length = r.f64(t, "lambda_length", false, 0.0);
            ts.connected = r.boolean(t, "connected", false);
            ts.max_activity = r.f64(t, "max_activity", false, 0.0);
            ts.lambda_activity = r.f64(t, "lambda_activity", false, 0.0);
            std::vector<double> ext = r.numbers(t, "external", false);
            if (!ext.empty() && ext.size() != 3) r.wrong("external", "an array of length 3");
            for (size_t k = 0; k < ext.size() && k < 3; ++k) ts.external[k] = ext[k];
            ts.lambda_nematic = r.f64(t, "lambda_nematic", false, 0.0);
            bp.types.push_back(ts);
        }
    if (const Json *fs = root.get("fields")) {
        if (fs->kind != Json::Array) r.wrong("fields", "a sequence");
        else
            for (const auto &f : fs->arr) {
                SpeciesSpec sp;
                sp.name = r.str(f, "name", true);
                sp.diffusion = r.f64(f, "diffusion", true, 0.0);
                sp.decay = r.f64(f, "decay", false, 0.0);
                sp.initial = r.f64(f, "initial", false, 0.0);
                sp.substeps = (int)r.uint(f, "substeps", false, 0);
                bp.fields.push_back(sp);
            }
    }
    if (const Json *ad = root.get("adhesion")) {
        if (ad->kind != Json::Null) {
            bp.has_adhesion = true;
            const Json *mol = ad->get("molecules");
            if (!mol) r.missing("molecules");
            else if (mol->kind != Json::Array) r.wrong("molecules", "a sequence");
            else
                for (const auto &m : mol->arr) bp.molecules.push_back(m.str);
            bp.binding = r.numbers(*ad, "binding", true);
        }
    }
    bp.contact = r.numbers(root, "contact", true);
    bp.temperature = r.f64(root, "temperature", true, 0.0);
    bp.neighbour_order = (uint32_t)r.uint(root, "neighbour_order", true, 0);
    bp.seed = r.uint(root, "seed", true, 0);
    bp.steps = r.uint(root, "steps", true, 0);
    bp.dump_every = r.uint(root, "dump_every", true, 0);
    if (const Json *nf = root.get("nematic_field")) {
        if (nf->kind == Json::String) {
            bp.has_nematic_field = true;
            bp.nematic_field = nf->str;
        }
    }
    const Json *ini = root.get("initial");
    if (!ini) r.missing("initial");
    else {
        bp.side = (uint32_t)r.uint(*ini, "side", true, 0);
        bp.nx = (uint32_t)r.uint(*ini, "nx", true, 0);
        bp.ny = (uint32_t)r.uint(*ini, "ny", true, 0);
        bp.nz = (uint32_t)r.uint(*ini, "nz", false, 1);
        std::map<std::string, double> fr = r.map(*ini, "fractions");
        bp.fractions.assign(fr.begin(), fr.end());
        if (const Json *lb = ini->get("labels")) {
            if (lb->kind == Json::String) {
                bp.has_labels = true;
                bp.labels = lb->str;
            }
        }
        if (ini->get("shuffle_seed")) {
            bp.has_shuffle = true;
            bp.shuffle_seed = r.uint(*ini, "shuffle_seed", false, 0);
        }
    }
    const Json *units = root.get("units");
    if (!units) r.missing("units");
    else {
        bp.micron_per_site = r.f64(*units, "micron_per_site", true, 0.0);
        bp.minute_per_step = r.f64(*units, "minute_per_step", true, 0.0);
    }
    if (!r.ok()) {
        *error = r.err;
        return false;
    }

    // the checks of Blueprint::from_json (blueprint.rs:386-446)
    for (const auto &f : bp.fractions) {
        bool known = false;
        for (const auto &t : bp.types) known |= t.name == f.first;
        if (!known) {
            *error = "the initial condition names the type " + f.first +
                     ", which the description does not state";
            return false;
        }
    }
    if (bp.has_adhesion) {
        size_t n = bp.molecules.size();
        if (bp.binding.size() != n * n) {
            *error = "the binding matrix is " + std::to_string(bp.binding.size()) + " entries for " +
                     std::to_string(n) + " molecules, want " + std::to_string(n * n);
            return false;
        }
        for (const auto &t : bp.types)
            for (const auto &p : t.presents)
                if (std::find(bp.molecules.begin(), bp.molecules.end(), p.first) == bp.molecules.end()) {
                    *error = "type " + t.name + " presents " + p.first +
                             ", which is not a molecule this description declares";
                    return false;
                }
    } else {
        for (const auto &t : bp.types)
            if (!t.presents.empty()) {
                *error = "type " + t.name +
                         " presents adhesion molecules, and the description declares none";
                return false;
            }
    }
    for (const auto &t : bp.types) {
        const std::pair<const char *, const std::map<std::string, double> *> lists[3] = {
            {"secretion", &t.secretion}, {"uptake", &t.uptake}, {"chemotaxis", &t.chemotaxis}};
        for (const auto &l : lists)
            for (const auto &kv : *l.second) {
                bool known = false;
                for (const auto &f : bp.fields) known |= f.name == kv.first;
                if (!known) {
                    *error = "type " + t.name + " states " + l.first + " of " + kv.first +
                             ", which is not a field this description declares";
                    return false;
                }
            }
    }
    if (bp.types.size() + 1 > GHIP_MAX_TYPES) {
        *error = "the hip engine takes at most 15 cell types beside the medium";
        return false;
    }
    if (bp.fields.size() > GHIP_MAX_SPECIES) {
        *error = "the hip engine takes at most 4 species";
        return false;
    }
    ghip_model m = bp.model();
    if (!validate_model(m, error)) return false;
    *out = bp;
    return true;
}

// effective_contact (blueprint.rs:200-239)
std::vector<double> Blueprint::effective_contact() const {
    std::vector<double> c = contact;
    if (!has_adhesion) return c;
    const size_t nt = types.size() + 1, n = molecules.size();
    std::vector<std::vector<double>> presented(nt, std::vector<double>(n, 0.0));
    for (size_t a = 1; a < nt; ++a)
        for (size_t i = 0; i < n; ++i) {
            auto it = types[a - 1].presents.find(molecules[i]);
            presented[a][i] = it == types[a - 1].presents.end() ? 0.0 : it->second;
        }
    for (size_t a = 0; a < nt; ++a)
        for (size_t b = 0; b < nt; ++b) {
            double bound = 0.0;
            for (size_t i = 0; i < n; ++i)
                for (size_t j = 0; j < n; ++j) bound += binding[i * n + j] * presented[a][i] * presented[b][j];
            if (a * nt + b < c.size()) c[a * nt + b] -= bound;
        }
    return c;
}

// initial_types (blueprint.rs:247-261), with the optional shuffle extension.
std::vector<uint8_t> Blueprint::initial_types(size_t cells) const {
    std::vector<uint8_t> out(cells, 1);
    size_t next = 0;
    for (size_t index = 0; index < types.size(); ++index) {
        double fraction = -1.0;
        for (const auto &f : fractions)
            if (f.first == types[index].name) fraction = f.second;
        if (fraction < 0.0) continue;
        size_t share = (size_t)std::llround(fraction * (double)cells);
        share = std::min(share, cells - next);
        for (size_t k = next; k < next + share; ++k) out[k] = (uint8_t)(index + 1);
        next += share;
    }
    if (has_shuffle && cells > 1) {
        uint32_t key = pass_key(shuffle_seed, 0, 0, 0x5EED5EEDu);
        for (size_t i = cells - 1; i > 0; --i) {
            uint32_t h = hash32(key ^ hash32((uint32_t)i));
            size_t j = (size_t)(((uint64_t)h * (uint64_t)(i + 1)) >> 32);
            std::swap(out[i], out[j]);
        }
    }
    return out;
}

// model (blueprint.rs:265-329)
ghip_model Blueprint::model() const {
    ghip_model m;
    std::memset(&m, 0, sizeof m);
    m.width = width;
    m.height = height;
    m.depth = depth;
    m.neighbour_order = neighbour_order;
    m.n_types = (uint32_t)types.size() + 1;
    m.n_species = (uint32_t)fields.size();
    m.temperature = temperature;
    m.seed = seed;
    std::vector<double> c = effective_contact();
    for (size_t k = 0; k < c.size() && k < (size_t)(GHIP_MAX_TYPES * GHIP_MAX_TYPES); ++k) m.contact[k] = c[k];
    for (size_t i = 0; i < types.size() && i + 1 < GHIP_MAX_TYPES; ++i) {
        const TypeSpec &t = types[i];
        ghip_cell_type &ct = m.types[i + 1];
        ct.target_volume = t.target_volume;
        ct.lambda_volume = t.lambda_volume;
        ct.target_surface = t.target_surface;
        ct.lambda_surface = t.lambda_surface;
        ct.division_volume = t.division_volume;
        ct.death_rate = t.death_rate;
        ct.target_length = t.target_length;
        ct.lambda_length = t.lambda_length;
        ct.max_activity = t.max_activity;
        ct.lambda_activity = t.lambda_activity;
        for (int a = 0; a < 3; ++a) ct.external[a] = t.external[a];
        ct.lambda_nematic = t.lambda_nematic;
        ct.connected = t.connected ? 1 : 0;
        for (size_t k = 0; k < fields.size() && k < GHIP_MAX_SPECIES; ++k) {
            auto rate = [&](const std::map<std::string, double> &mp) {
                auto it = mp.find(fields[k].name);
                return it == mp.end() ? 0.0 : it->second;
            };
            m.secretion[i + 1][k] = rate(t.secretion);
            m.uptake[i + 1][k] = rate(t.uptake);
            m.chemotaxis[i + 1][k] = rate(t.chemotaxis);
        }
    }
    for (size_t k = 0; k < fields.size() && k < GHIP_MAX_SPECIES; ++k) {
        m.species[k].diffusion = fields[k].diffusion;
        m.species[k].decay = fields[k].decay;
        m.species[k].initial = fields[k].initial;
        m.species[k].substeps = fields[k].substeps;
    }
    return m;
}

// Model::validate (model.rs:160-232); the species-table checks hold by
// construction of the flat model.
bool validate_model(const ghip_model &m, std::string *error) {
    const uint32_t n = m.n_types;
    if (n == 0) {
        *error = "a model needs at least the medium type";
        return false;
    }
    for (uint32_t a = 0; a < n; ++a)
        for (uint32_t b = 0; b < n; ++b)
            if (std::fabs(m.contact[a * n + b] - m.contact[b * n + a]) > 1e-12) {
                *error = "contact matrix is asymmetric at (" + std::to_string(a) + ", " +
                         std::to_string(b) + ")";
                return false;
            }
    if (m.width == 0 || m.height == 0 || m.depth == 0) {
        *error = "the lattice has a zero dimension";
        return false;
    }
    if (m.neighbour_order < 1 || m.neighbour_order > 3) {
        *error = "neighbour order " + std::to_string(m.neighbour_order) +
                 " is outside the range one to three";
        return false;
    }
    if (!(m.temperature > 0.0)) {
        *error = "the temperature must be positive";
        return false;
    }
    return true;
}

// Lattice::tile_grid (lattice.rs:169-196)
uint32_t tile_grid(const ghip_model &m, uint32_t side, uint32_t nx, uint32_t ny, uint32_t nz,
                   std::vector<uint32_t> *labels) {
    const size_t w = m.width, h = m.height, d = m.depth;
    labels->assign(w * h * d, 0u);
    if (side == 0) return 0;
    nx = std::min<uint32_t>(nx, (uint32_t)(w / side));
    ny = std::min<uint32_t>(ny, (uint32_t)(h / side));
    nz = d == 1 ? 1 : std::min<uint32_t>(nz, (uint32_t)(d / side));
    const size_t side_z = d == 1 ? 1 : side;
    uint32_t next = 1;
    for (uint32_t bz = 0; bz < nz; ++bz)
        for (uint32_t by = 0; by < ny; ++by)
            for (uint32_t bx = 0; bx < nx; ++bx) {
                for (size_t z = bz * side_z; z < (bz + 1) * side_z; ++z)
                    for (size_t y = by * side; y < (by + 1) * side; ++y)
                        for (size_t x = bx * side; x < (bx + 1) * side; ++x) (*labels)[(z * h + y) * w + x] = next;
                ++next;
            }
    return next - 1;
}

// ---------------------------------------------------------------- --check

namespace {
std::string num(double v) {
    char buf[40];
    if (v == std::floor(v) && std::fabs(v) < 1e15) std::snprintf(buf, sizeof buf, "%.1f", v);
    else std::snprintf(buf, sizeof buf, "%.17g", v);
    return buf;
}
std::string quoted(const std::string &s) {
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') o.push_back('\\');
        o.push_back(c);
    }
    return o + "\"";
}
template <class F>
std::string list(size_t n, F f) {
    std::string o = "[";
    for (size_t i = 0; i < n; ++i) o += (i ? ", " : "") + f(i);
    return o + "]";
}
}  // namespace

// The keys of glazier/src/main.rs:62-96 (serde_json prints them sorted).
std::string check_json(const Blueprint &bp, bool with_depth) {
    ghip_model m = bp.model();
    const size_t nt = m.n_types, ns = m.n_species;
    std::vector<std::pair<std::string, std::string>> kv;
    auto per_type = [&](double ghip_cell_type::*f) {
        return list(nt, [&](size_t i) { return num(m.types[i].*f); });
    };
    kv.push_back({"chemotaxis", list(nt, [&](size_t t) { return list(ns, [&](size_t k) { return num(m.chemotaxis[t][k]); }); })});
    kv.push_back({"connected", list(nt, [&](size_t i) { return std::string(m.types[i].connected ? "true" : "false"); })});
    kv.push_back({"contact", list(nt * nt, [&](size_t i) { return num(m.contact[i]); })});
    kv.push_back({"death_rate", per_type(&ghip_cell_type::death_rate)});
    kv.push_back({"decay", list(ns, [&](size_t k) { return num(m.species[k].decay); })});
    if (with_depth) kv.push_back({"depth", std::to_string(m.depth)});
    kv.push_back({"diffusion", list(ns, [&](size_t k) { return num(m.species[k].diffusion); })});
    kv.push_back({"division_volume", per_type(&ghip_cell_type::division_volume)});
    kv.push_back({"external", list(nt, [&](size_t i) { return list(3, [&](size_t a) { return num(m.types[i].external[a]); }); })});
    kv.push_back({"height", std::to_string(m.height)});
    kv.push_back({"initial", "[" + std::to_string(bp.side) + ", " + std::to_string(bp.nx) + ", " +
                                 std::to_string(bp.ny) + ", " + std::to_string(bp.nz) + "]"});
    kv.push_back({"initial_labels", bp.has_labels ? quoted(bp.labels) : "null"});
    kv.push_back({"lambda_activity", per_type(&ghip_cell_type::lambda_activity)});
    kv.push_back({"lambda_length", per_type(&ghip_cell_type::lambda_length)});
    kv.push_back({"lambda_nematic", per_type(&ghip_cell_type::lambda_nematic)});
    kv.push_back({"lambda_surface", per_type(&ghip_cell_type::lambda_surface)});
    kv.push_back({"lambda_volume", per_type(&ghip_cell_type::lambda_volume)});
    kv.push_back({"max_activity", per_type(&ghip_cell_type::max_activity)});
    kv.push_back({"micron_per_site", num(bp.micron_per_site)});
    kv.push_back({"minute_per_step", num(bp.minute_per_step)});
    kv.push_back({"n_types", std::to_string(nt)});
    kv.push_back({"name", quoted(bp.name)});
    kv.push_back({"neighbour_order", std::to_string(m.neighbour_order)});
    kv.push_back({"nematic_field", bp.has_nematic_field ? quoted(bp.nematic_field) : "null"});
    kv.push_back({"secretion", list(nt, [&](size_t t) { return list(ns, [&](size_t k) { return num(m.secretion[t][k]); }); })});
    kv.push_back({"seed", std::to_string(m.seed)});
    kv.push_back({"species", list(ns, [&](size_t k) { return quoted(bp.fields[k].name); })});
    kv.push_back({"steps", std::to_string(bp.steps)});
    kv.push_back({"target_length", per_type(&ghip_cell_type::target_length)});
    kv.push_back({"target_surface", per_type(&ghip_cell_type::target_surface)});
    kv.push_back({"target_volume", per_type(&ghip_cell_type::target_volume)});
    kv.push_back({"temperature", num(m.temperature)});
    kv.push_back({"uptake", list(nt, [&](size_t t) { return list(ns, [&](size_t k) { return num(m.uptake[t][k]); }); })});
    kv.push_back({"width", std::to_string(m.width)});
    std::string o = "{\n";
    for (size_t i = 0; i < kv.size(); ++i) o += "  " + quoted(kv[i].first) + ": " + kv[i].second + (i + 1 < kv.size() ? ",\n" : "\n");
    return o + "}";
}

// ---------------------------------------------------------------- .npy

namespace {
bool npy_write(const std::string &path, const void *data, size_t elem, const char *descr, uint32_t w,
               uint32_t h, uint32_t d, std::string *error) {
    std::string shape = d > 1 ? "(" + std::to_string(d) + ", " + std::to_string(h) + ", " + std::to_string(w) + ")"
                              : "(" + std::to_string(h) + ", " + std::to_string(w) + ")";
    std::string header = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': " + shape + ", }";
    size_t unpadded = 10 + header.size() + 1;
    size_t padding = (64 - unpadded % 64) % 64;
    header += std::string(padding, ' ') + "\n";
    std::ofstream f(path, std::ios::binary);
    if (!f) {
        *error = "cannot create " + path;
        return false;
    }
    f.write("\x93NUMPY", 6);
    const char ver[2] = {1, 0};
    f.write(ver, 2);
    uint16_t len = (uint16_t)header.size();
    f.write(reinterpret_cast<const char *>(&len), 2);
    f.write(header.data(), (std::streamsize)header.size());
    f.write(reinterpret_cast<const char *>(data), (std::streamsize)(elem * (size_t)w * h * d));
    return (bool)f;
}
}  // namespace

// npy::write_u32 / write_f64 (npy.rs:17-94)
bool npy_write_u32(const std::string &path, const uint32_t *data, uint32_t w, uint32_t h, uint32_t d,
                   std::string *error) {
    return npy_write(path, data, 4, "<u4", w, h, d, error);
}
bool npy_write_f64(const std::string &path, const double *data, uint32_t w, uint32_t h, uint32_t d,
                   std::string *error) {
    return npy_write(path, data, 8, "<f8", w, h, d, error);
}

namespace {
// The parts of a .npy file the readers need (npy.rs:96-144): dtype string,
// shape, and the data that follows the header.
struct NpyFile {
    std::string bytes;
    std::string descr;
    std::vector<size_t> shape;
    size_t count = 1;
    size_t offset = 0;
};

bool npy_open(const std::string &path, NpyFile *f, std::string *error) {
    std::ifstream in(path, std::ios::binary);
    if (!in) {
        *error = "cannot read " + path;
        return false;
    }
    f->bytes.assign((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    const std::string &bytes = f->bytes;
    if (bytes.size() < 10 || bytes.compare(0, 6, "\x93NUMPY") != 0) {
        *error = "not a .npy file";
        return false;
    }
    size_t len, start;
    if (bytes[6] == 1) {
        len = (uint8_t)bytes[8] | ((size_t)(uint8_t)bytes[9] << 8);
        start = 10;
    } else {
        if (bytes.size() < 12) {
            *error = "not a .npy file";
            return false;
        }
        len = (uint8_t)bytes[8] | ((size_t)(uint8_t)bytes[9] << 8) | ((size_t)(uint8_t)bytes[10] << 16) |
              ((size_t)(uint8_t)bytes[11] << 24);
        start = 12;
    }
    if (start + len > bytes.size()) {
        *error = "the .npy header runs past the file";
        return false;
    }
    std::string header = bytes.substr(start, len);
    if (header.find("'fortran_order': True") != std::string::npos) {
        *error = "Fortran-order arrays are not read";
        return false;
    }
    size_t dp = header.find("'descr':");
    size_t q1 = header.find('\'', dp + 8), q2 = header.find('\'', q1 + 1);
    size_t sp = header.find('(', header.find("'shape':")), se = header.find(')', sp);
    if (dp == std::string::npos || q2 == std::string::npos || sp == std::string::npos ||
        se == std::string::npos) {
        *error = "the .npy header has no dtype or shape";
        return false;
    }
    f->descr = header.substr(q1 + 1, q2 - q1 - 1);
    std::stringstream ss(header.substr(sp + 1, se - sp - 1));
    std::string tok;
    f->count = 1;
    while (std::getline(ss, tok, ',')) {
        size_t a = tok.find_first_not_of(' ');
        if (a == std::string::npos) continue;
        f->shape.push_back((size_t)std::stoull(tok.substr(a)));
        f->count *= f->shape.back();
    }
    f->offset = start + len;
    return true;
}
}  // namespace

// npy::read_labels (npy.rs:146-176): <u4, <i4 (non-negative), <i8/<u8.
bool npy_read_labels(const std::string &path, std::vector<uint32_t> *out, std::string *error) {
    NpyFile f;
    if (!npy_open(path, &f, error)) return false;
    const size_t count = f.count;
    const char *data = f.bytes.data() + f.offset;
    const size_t avail = f.bytes.size() - f.offset;
    const std::string &descr = f.descr;
    out->resize(count);
    if (descr == "<u4" || descr == "<i4") {
        if (avail < count * 4) {
            *error = "the .npy file is shorter than its shape";
            return false;
        }
        for (size_t i = 0; i < count; ++i) {
            int32_t v;
            std::memcpy(&v, data + 4 * i, 4);
            if (descr == "<i4" && v < 0) {
                *error = "negative label " + std::to_string(v);
                return false;
            }
            (*out)[i] = (uint32_t)v;
        }
    } else if (descr == "<i8" || descr == "<u8") {
        if (avail < count * 8) {
            *error = "the .npy file is shorter than its shape";
            return false;
        }
        for (size_t i = 0; i < count; ++i) {
            int64_t v;
            std::memcpy(&v, data + 8 * i, 8);
            if (v < 0 || v > 0xFFFFFFFFLL) {
                *error = "label " + std::to_string(v) + " is outside u32";
                return false;
            }
            (*out)[i] = (uint32_t)v;
        }
    } else {
        *error = "labels of dtype " + descr + " are not read";
        return false;
    }
    return true;
}

// npy::read_pairs (npy.rs:188-214): a last axis of 2, <f8 or <f4; the values
// in file order, two per site.
bool npy_read_pairs(const std::string &path, std::vector<double> *out, std::string *error) {
    NpyFile f;
    if (!npy_open(path, &f, error)) return false;
    if (f.shape.empty() || f.shape.back() != 2) {
        *error = "the last axis is " + (f.shape.empty() ? std::string("missing") : std::to_string(f.shape.back())) +
                 ", want 2";
        return false;
    }
    const char *data = f.bytes.data() + f.offset;
    const size_t avail = f.bytes.size() - f.offset;
    const size_t elem = f.descr == "<f8" ? 8 : f.descr == "<f4" ? 4 : 0;
    if (!elem) {
        *error = "a field of dtype " + f.descr + " is not read";
        return false;
    }
    if (avail < f.count * elem) {
        *error = "the .npy file is shorter than its shape";
        return false;
    }
    out->resize(f.count);
    for (size_t i = 0; i < f.count; ++i) {
        if (elem == 8) {
            std::memcpy(&(*out)[i], data + 8 * i, 8);
        } else {
            float v;
            std::memcpy(&v, data + 4 * i, 4);
            (*out)[i] = (double)v;
        }
    }
    return true;
}

}  // namespace ghip
