// Formats on either side of the transport step (host code, no device work):
//  * the H/C handoff written after a preheat and read before an evolution:
//    cellH1.dat / cellC1.dat (text matrices, one row per j, x-fast, 17
//    significant digits) and the transport_state.json manifest
//    (src/couple/initialization.cpp:642-686, 823-841; reader
//    src/couple/evolve.cpp:54-125);
//  * the edge Gauss-point velocity samples of the flow solve (FlowPointSample,
//    src/couple/visualization.h:8-12, filled by SampleInitialFlow,
//    src/couple/flow_sampling.cpp:62-121) turned into the face-velocity arrays
//    of mp_transport_set_face_velocity, as the preheat builder does
//    (src/couple/preheat.cpp:83-90: the last sample of an (edge, q) wins).
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "mp_host.h"

using namespace mp;

namespace {

// ---------------------------------------------------------------- JSON
// A small reader for the manifest: objects, arrays, strings, numbers, true,
// false, null.  Numbers keep their text so 64-bit seeds survive exactly.
struct Json {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0;
    std::string text;                       // string value, or number literal
    std::vector<Json> items;
    std::vector<std::pair<std::string, Json>> fields;

    const Json* get(const char* key) const
    {
        if (kind != Object) return nullptr;
        for (const auto& f : fields)
            if (f.first == key) return &f.second;
        return nullptr;
    }
};

struct JsonParser {
    const std::string& s;
    size_t i = 0;
    std::string error;
    explicit JsonParser(const std::string& src) : s(src) {}

    void ws()
    {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\r' || s[i] == '\t')) ++i;
    }
    bool fail(const char* what)
    {
        if (error.empty()) error = std::string(what) + " at offset " + std::to_string(i);
        return false;
    }
    bool string(std::string& out)
    {
        if (i >= s.size() || s[i] != '"') return fail("expected a string");
        ++i;
        out.clear();
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= s.size()) return fail("unterminated escape");
                const char e = s[i++];
                switch (e) {
                case '"': c = '"'; break;
                case '\\': c = '\\'; break;
                case '/': c = '/'; break;
                case 'b': c = '\b'; break;
                case 'f': c = '\f'; break;
                case 'n': c = '\n'; break;
                case 'r': c = '\r'; break;
                case 't': c = '\t'; break;
                case 'u': {
                    if (i + 4 > s.size()) return fail("short unicode escape");
                    const unsigned v = unsigned(std::strtoul(s.substr(i, 4).c_str(), nullptr, 16));
                    i += 4;
                    if (v < 0x80) { c = char(v); break; }
                    out += '?';   // outside the manifest's ASCII vocabulary
                    continue;
                }
                default: return fail("unknown escape");
                }
            }
            out += c;
        }
        if (i >= s.size()) return fail("unterminated string");
        ++i;
        return true;
    }
    bool value(Json& v, int depth = 0)
    {
        if (depth > 32) return fail("nesting too deep");
        ws();
        if (i >= s.size()) return fail("unexpected end");
        const char c = s[i];
        if (c == '{') {
            ++i;
            v.kind = Json::Object;
            ws();
            if (i < s.size() && s[i] == '}') { ++i; return true; }
            for (;;) {
                ws();
                std::string key;
                if (!string(key)) return false;
                ws();
                if (i >= s.size() || s[i] != ':') return fail("expected ':'");
                ++i;
                Json item;
                if (!value(item, depth + 1)) return false;
                v.fields.emplace_back(key, std::move(item));
                ws();
                if (i < s.size() && s[i] == ',') { ++i; continue; }
                if (i < s.size() && s[i] == '}') { ++i; return true; }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            ++i;
            v.kind = Json::Array;
            ws();
            if (i < s.size() && s[i] == ']') { ++i; return true; }
            for (;;) {
                Json item;
                if (!value(item, depth + 1)) return false;
                v.items.push_back(std::move(item));
                ws();
                if (i < s.size() && s[i] == ',') { ++i; continue; }
                if (i < s.size() && s[i] == ']') { ++i; return true; }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = Json::String;
            return string(v.text);
        }
        if (s.compare(i, 4, "true") == 0) { i += 4; v.kind = Json::Bool; v.b = true; return true; }
        if (s.compare(i, 5, "false") == 0) { i += 5; v.kind = Json::Bool; v.b = false; return true; }
        if (s.compare(i, 4, "null") == 0) { i += 4; v.kind = Json::Null; return true; }
        // number (also NaN / Infinity as Python's json writes them)
        const size_t start = i;
        while (i < s.size() && (std::strchr("+-0123456789.eE", s[i]) || std::isalpha(static_cast<unsigned char>(s[i])))) ++i;
        if (i == start) return fail("unexpected character");
        v.kind = Json::Number;
        v.text = s.substr(start, i - start);
        char* end = nullptr;
        errno = 0;
        v.num = std::strtod(v.text.c_str(), &end);
        if (end != v.text.c_str() + v.text.size()) return fail("malformed number");
        return true;
    }
    bool document(Json& v)
    {
        if (!value(v)) return false;
        ws();
        if (i != s.size()) return fail("trailing characters");
        return true;
    }
};

bool read_file(const std::string& path, std::string& out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::stringstream ss;
    ss << f.rdbuf();
    out = ss.str();
    return true;
}

bool close_rel(double a, double b)
{
    // the reader's tolerance for domain and scales (src/couple/evolve.cpp:70)
    return std::isfinite(a) && std::fabs(a - b) <= 1e-12 * std::fmax(1.0, std::fmax(std::fabs(a), std::fabs(b)));
}

// Write text to path through a temporary file and a rename, so a reader never
// sees a half-written file (src/couple/initialization.cpp:662-683).
mp_status write_atomic(const std::string& path, const std::string& text)
{
    const std::string tmp = path + ".tmp";
    {
        std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
        if (!f) return fail(MP_ERR_ARG, "Could not write %s", tmp.c_str());
        f << text;
        f.close();
        if (!f) return fail(MP_ERR_ARG, "Could not write %s", tmp.c_str());
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0)
        return fail(MP_ERR_ARG, "Could not rename %s: %s", tmp.c_str(), std::strerror(errno));
    return MP_OK;
}

std::string fmt17(double v)
{
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    return buf;
}

std::string matrix_text(const double* u, int nx, int ny)
{
    std::string t;
    t.reserve(size_t(nx) * ny * 24);
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
            if (i) t += ' ';
            t += fmt17(u[size_t(j) * nx + i]);
        }
        t += '\n';
    }
    return t;
}

// Whitespace-separated doubles; exactly n of them (src/couple/evolve.cpp:79-85).
mp_status read_matrix(const std::string& path, size_t n, double cmax, bool composition, double* out)
{
    std::string text;
    if (!read_file(path, text)) return fail(MP_ERR_ARG, "Cannot open preheat field %s", path.c_str());
    const char* p = text.c_str();
    const char* end = p + text.size();
    size_t k = 0;
    for (;;) {
        while (p < end && std::isspace(static_cast<unsigned char>(*p))) ++p;
        if (p >= end) break;
        char* q = nullptr;
        const double v = std::strtod(p, &q);
        if (q == p) return fail(MP_ERR_ARG, "Incomplete preheat cell-average matrix (unreadable entry %zu)", k);
        if (k >= n) return fail(MP_ERR_ARG, "Extra entries in preheat cell-average matrix");
        if (!std::isfinite(v) || v < 0.0 || (composition && v > cmax))
            return fail(MP_ERR_ARG, "Inadmissible preheat cell average %g at entry %zu of %s", v, k, path.c_str());
        out[k++] = v;
        p = q;
    }
    if (k < n) return fail(MP_ERR_ARG, "Incomplete preheat cell-average matrix (%zu of %zu entries)", k, n);
    return MP_OK;
}

const char* str_of(const Json* v) { return v && v->kind == Json::String ? v->text.c_str() : nullptr; }
bool num_of(const Json* v, double& out)
{
    if (!v || v->kind != Json::Number) return false;
    out = v->num;
    return true;
}

}  // namespace

// ===================================================================== C ABI

extern "C" {

mp_status mp_handoff_write(const char* directory, const mp_handoff_meta* meta, const double* H,
                           const double* C)
{
    if (!directory || !meta || !H || !C) return fail(MP_ERR_ARG, "null argument");
    MP_GUARD({
        const int nx = meta->nx, ny = meta->ny;
        if (nx < 1 || ny < 1) return fail(MP_ERR_ARG, "mesh dimensions must be positive");
        const std::string dir(directory);
        MP_CALL(write_atomic(dir + "/cellH1.dat", matrix_text(H, nx, ny)));
        MP_CALL(write_atomic(dir + "/cellC1.dat", matrix_text(C, nx, ny)));
        // manifest in the layout of src/couple/initialization.cpp:823-839
        std::string fam(meta->family, strnlen(meta->family, sizeof meta->family));
        std::ostringstream m;
        m << "{\n\"schema_version\":1,\"status\":\"complete\","
          << "\"representation\":\"cell averages\",\"ordering\":\"j*nx+i\",\n"
          << "\"H\":{\"file\":\"cellH1.dat\",\"units\":\"h/(cp*dT)\"},"
          << "\"C\":{\"file\":\"cellC1.dat\",\"units\":\"fraction\"},\n"
          << "\"time\":" << fmt17(meta->time) << ",\"time_s\":" << fmt17(meta->time * meta->time_scale_s) << ",\n"
          << "\"accepted_steps\":" << (long long)meta->accepted_steps << ",\n"
          << "\"mesh\":{\"family\":\"" << fam << "\",\"nx\":" << nx << ",\"ny\":" << ny
          << ",\"domain\":[" << fmt17(meta->domain[0]) << ',' << fmt17(meta->domain[1]) << ','
          << fmt17(meta->domain[2]) << ',' << fmt17(meta->domain[3]) << ']'
          << ",\"perturbation\":" << fmt17(meta->perturbation) << ",\"seed\":" << (unsigned long long)meta->seed << "},\n"
          << "\"scales\":{\"enthalpy_Jkg\":" << fmt17(meta->enthalpy_Jkg) << ",\"length_m\":"
          << fmt17(meta->length_m) << ",\"time_s\":" << fmt17(meta->time_scale_s) << "}\n}\n";
        return write_atomic(dir + "/transport_state.json", m.str());
    })
}

mp_status mp_handoff_read(const char* directory, const mp_handoff_meta* expect, double c_max,
                          double* H, double* C, mp_handoff_meta* found)
{
    if (!directory || !expect || !H || !C) return fail(MP_ERR_ARG, "null argument");
    MP_GUARD({
        const std::string dir(directory);
        std::string text;
        if (!read_file(dir + "/transport_state.json", text))
            return fail(MP_ERR_ARG, "Cannot open %s/transport_state.json", directory);
        Json m;
        JsonParser parser(text);
        if (!parser.document(m) || m.kind != Json::Object)
            return fail(MP_ERR_ARG, "Malformed preheat manifest: %s", parser.error.c_str());
        // checks of src/couple/evolve.cpp:62-78, in that order
        double schema = 0;
        const char* status = str_of(m.get("status"));
        if (!num_of(m.get("schema_version"), schema) || schema != 1.0 || !status || std::strcmp(status, "complete"))
            return fail(MP_ERR_ARG, "Preheat manifest must be complete schema 1");
        const char* rep = str_of(m.get("representation"));
        const char* ord = str_of(m.get("ordering"));
        if (!rep || !ord || std::strcmp(rep, "cell averages") || std::strcmp(ord, "j*nx+i"))
            return fail(MP_ERR_ARG, "Unsupported preheat representation/order");
        const Json* mh = m.get("H");
        const Json* mc = m.get("C");
        const char* hu = str_of(mh ? mh->get("units") : nullptr);
        const char* cu = str_of(mc ? mc->get("units") : nullptr);
        if (!hu || !cu || std::strcmp(hu, "h/(cp*dT)") || std::strcmp(cu, "fraction"))
            return fail(MP_ERR_ARG, "Preheat units do not match");
        const char* hf = str_of(mh->get("file"));
        const char* cf = str_of(mc->get("file"));
        if (!hf || !cf || std::strcmp(hf, "cellH1.dat") || std::strcmp(cf, "cellC1.dat"))
            return fail(MP_ERR_ARG, "Unsupported preheat filenames");
        const Json* mesh = m.get("mesh");
        const char* fam = str_of(mesh ? mesh->get("family") : nullptr);
        double nx = -1, ny = -1;
        const std::string want(expect->family, strnlen(expect->family, sizeof expect->family));
        if (!fam || want != fam || !num_of(mesh->get("nx"), nx) || !num_of(mesh->get("ny"), ny) ||
            nx != expect->nx || ny != expect->ny)
            return fail(MP_ERR_ARG, "Preheat mesh does not match evolution mesh");
        const Json* dom = mesh->get("domain");
        if (!dom || dom->kind != Json::Array || dom->items.size() != 4)
            return fail(MP_ERR_ARG, "Preheat domain does not match");
        for (int k = 0; k < 4; ++k) {
            double v;
            if (!num_of(&dom->items[k], v) || !close_rel(v, expect->domain[k]))
                return fail(MP_ERR_ARG, "Preheat domain does not match");
        }
        double pert = 0;
        unsigned long long seed = 0;
        const Json* ps = mesh->get("seed");
        const bool have_pert = num_of(mesh->get("perturbation"), pert);
        if (ps && ps->kind == Json::Number) seed = std::strtoull(ps->text.c_str(), nullptr, 10);
        if (want != "rectangular" &&
            (!have_pert || !close_rel(pert, expect->perturbation) || !ps || seed != expect->seed))
            return fail(MP_ERR_ARG, "Preheat mesh perturbation does not match");
        const Json* sc = m.get("scales");
        double h0 = 0, l0 = 0, t0 = 0;
        if (!sc || !num_of(sc->get("enthalpy_Jkg"), h0) || !num_of(sc->get("length_m"), l0) ||
            !num_of(sc->get("time_s"), t0) || !close_rel(h0, expect->enthalpy_Jkg) ||
            !close_rel(l0, expect->length_m) || !close_rel(t0, expect->time_scale_s))
            return fail(MP_ERR_ARG, "Preheat scales do not match; explicit physical conversion is required");
        double time = 0;
        if (!num_of(m.get("time"), time) || !std::isfinite(time)) return fail(MP_ERR_FP, "Nonfinite source time");
        const size_t n = size_t(expect->nx) * size_t(expect->ny);
        std::vector<double> h(n), c(n);
        MP_CALL(read_matrix(dir + "/cellH1.dat", n, c_max, false, h.data()));
        MP_CALL(read_matrix(dir + "/cellC1.dat", n, c_max, true, c.data()));
        std::memcpy(H, h.data(), n * sizeof(double));
        std::memcpy(C, c.data(), n * sizeof(double));
        if (found) {
            mp_handoff_meta f = *expect;
            f.time = time;
            f.perturbation = have_pert ? pert : expect->perturbation;
            f.seed = ps ? seed : expect->seed;
            double steps = -1;
            f.accepted_steps = num_of(m.get("accepted_steps"), steps) ? int64_t(steps) : -1;
            *found = f;
        }
        return MP_OK;
    })
}

mp_status mp_face_velocity_from_samples(int32_t nx, int32_t ny, int32_t q, int64_t nsamples,
                                        const mp_edge_sample* samples, double* vnx, double* vny)
{
    if (nsamples > 0 && !samples) return fail(MP_ERR_ARG, "null sample array");
    if (!vnx || !vny) return fail(MP_ERR_ARG, "null output array");
    if (nx < 1 || ny < 1 || q < 1 || q > kMaxQ) return fail(MP_ERR_ARG, "bad mesh or point count");
    MP_GUARD({
        const int64_t along_i = int64_t(nx) * (ny + 1);            // horizontal edges first
        const int64_t edges = along_i + int64_t(nx + 1) * ny;
        const size_t nxv = size_t(q) * ny * (nx + 1), nyv = size_t(q) * (ny + 1) * nx;
        std::vector<unsigned char> seen(nxv + nyv, 0);
        std::vector<double> tx(nxv), ty(nyv);   // outputs are assigned on success only
        for (int64_t k = 0; k < nsamples; ++k) {
            const mp_edge_sample& s = samples[k];
            if (s.edge_id < 0 || s.edge_id >= edges)
                return fail(MP_ERR_ARG, "Global edge ID lies outside the mesh (sample %lld)", (long long)k);
            if (s.q < 0 || s.q >= q) return fail(MP_ERR_ARG, "edge point index outside the rule (sample %lld)", (long long)k);
            if (!std::isfinite(s.vx) || !std::isfinite(s.vy))
                return fail(MP_ERR_FP, "Nonfinite reconstructed velocity (sample %lld)", (long long)k);
            size_t at;
            if (s.edge_id < along_i) {
                // horizontal edge (i, j): id = j*nx + i, canonical direction +x
                const int64_t j = s.edge_id / nx, i = s.edge_id % nx;
                at = (size_t(s.q) * (ny + 1) + size_t(j)) * nx + size_t(i);
                ty[at] = s.vy;
                seen[nxv + at] = 1;
            } else {
                // vertical edge (i, j): id = nx*(ny+1) + j*(nx+1) + i, canonical direction +y
                const int64_t e = s.edge_id - along_i, j = e / (nx + 1), i = e % (nx + 1);
                at = (size_t(s.q) * ny + size_t(j)) * (nx + 1) + size_t(i);
                tx[at] = s.vx;
                seen[at] = 1;
            }
        }
        for (size_t k = 0; k < seen.size(); ++k)
            if (!seen[k]) return fail(MP_ERR_ARG, "no velocity sample for face point %zu", k);
        std::memcpy(vnx, tx.data(), nxv * sizeof(double));
        std::memcpy(vny, ty.data(), nyv * sizeof(double));
        return MP_OK;
    })
}

}  // extern "C"
