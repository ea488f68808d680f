// The reference's on-disk state of a species and a deck, byte for byte, and
// reading it back (the checkpoint of device-resident and served decks).
//
// Under <sim_dir>/simulation_output/<species>/ (AMRStructure_io.cpp:3-101,
// AMRStructure_restart.cpp:33-98):
//   xs/xs_<iter>, ps/ps_<iter>, fs/fs_<iter>     raw doubles, one per point
//   qws/qws_<iter>, es/es_<iter>                 raw doubles (the files the
//                                                reference's diagnostics read)
//   panels/leaf_point_inds_<iter>                9 int32 per leaf, leaves in
//                                                depth-first order
//   panels/full_tree_<iter>                      int64 panel count, then one
//                                                80-byte record per panel
// and under <sim_dir>/simulation_output/ the deck's reduced set
// (AMRSimulation_field.cpp:81-111): reduce_xs/xs_<iter>, reduced_es/es_<iter>.
// Host code only; the caller brings host arrays. Files are written whole
// (a directory that is missing is created).
#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fs_internal.h"
#include "fs_mesh.h"

namespace fs {
namespace io {
namespace {

// One panel of panels/full_tree_<iter> (AMRStructure_restart.cpp:33-52):
// int32 fields, then the five flags as int8 and three bytes of padding.
struct PanelDump {
    int32_t panel_ind, level, parent_ind, which_child;
    int32_t point_inds[9];
    int32_t left, top, right, bottom;
    int32_t child_inds_start;
    int8_t left_bdry, right_bdry, refined_xp, refined_p, needs_refinement;
    int8_t pad[3];
};
static_assert(sizeof(PanelDump) == 80, "the reference's record is 80 bytes");

int make_dirs(const std::string& path) {
    std::string cur;
    size_t pos = 0;
    while (pos != std::string::npos) {
        pos = path.find('/', pos + 1);
        cur = path.substr(0, pos);
        if (cur.empty()) continue;
        if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
            return fail(FS_E_INVALID, "cannot create directory %s: %s", cur.c_str(),
                        std::strerror(errno));
    }
    return FS_OK;
}

int write_file(const std::string& path, const void* data, size_t bytes) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return fail(FS_E_INVALID, "cannot open %s for writing", path.c_str());
    const size_t w = bytes ? std::fwrite(data, 1, bytes, f) : 0;
    const int rc = std::fclose(f);
    if (w != bytes || rc != 0) return fail(FS_E_INVALID, "error writing %s", path.c_str());
    return FS_OK;
}

int read_file(const std::string& path, std::vector<char>& out) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return fail(FS_E_INVALID, "cannot open %s", path.c_str());
    std::fseek(f, 0, SEEK_END);
    const long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out.resize(size > 0 ? (size_t)size : 0);
    const size_t r = out.empty() ? 0 : std::fread(out.data(), 1, out.size(), f);
    std::fclose(f);
    if (size < 0 || r != out.size()) return fail(FS_E_INVALID, "error reading %s", path.c_str());
    return FS_OK;
}

std::string species_dir(const char* sim_dir, const char* name) {
    std::string d = sim_dir ? sim_dir : "";
    if (!d.empty() && d.back() != '/') d += '/';
    return d + "simulation_output/" + name + "/";
}

// Leaves in the depth-first order of set_leaves_weights (children in index
// order), from the panel records.
std::vector<int32_t> leaves_of(const int32_t* rec, int64_t n_panels) {
    std::vector<int32_t> leaves;
    if (n_panels <= 0) return leaves;
    std::vector<int32_t> stack(1, 0);
    while (!stack.empty()) {
        const int32_t p = stack.back();
        stack.pop_back();
        const int32_t* r = rec + (size_t)p * mesh::kInts;
        const int nch = (r[mesh::FLAGS] & mesh::REFINED_P) ? 2
                        : (r[mesh::FLAGS] & mesh::REFINED_XP) ? 4 : 0;
        if (nch == 0) {
            leaves.push_back(p);
            continue;
        }
        for (int c = nch - 1; c >= 0; --c) stack.push_back(r[mesh::CHILD] + c);
    }
    return leaves;
}

}  // namespace

int write_species(const char* sim_dir, const char* name, int iter, int64_t n, const double* xs,
                  const double* ps, const double* fs, const double* q_ws, const double* es,
                  int64_t n_panels, const int32_t* rec) {
    if (!name || !*name || iter < 0 || n < 0 || n_panels < 0 || (n > 0 && (!xs || !ps || !fs)) ||
        (n_panels > 0 && !rec))
        return fail(FS_E_INVALID, "write_species_output: bad argument");
    const std::string dir = species_dir(sim_dir, name), it = std::to_string(iter);
    const size_t nb = sizeof(double) * (size_t)n;
    for (const char* sub : {"xs", "ps", "fs", "qws", "es", "panels"})
        FS_TRY(make_dirs(dir + sub));
    FS_TRY(write_file(dir + "xs/xs_" + it, xs, nb));
    FS_TRY(write_file(dir + "ps/ps_" + it, ps, nb));
    FS_TRY(write_file(dir + "fs/fs_" + it, fs, nb));
    if (q_ws) FS_TRY(write_file(dir + "qws/qws_" + it, q_ws, nb));
    if (es) FS_TRY(write_file(dir + "es/es_" + it, es, nb));
    // leaves' point indices (AMRStructure_io.cpp:73-101)
    const std::vector<int32_t> leaves = leaves_of(rec, n_panels);
    std::vector<int32_t> lp;
    lp.reserve(leaves.size() * 9);
    for (int32_t p : leaves)
        for (int k = 0; k < 9; ++k) lp.push_back(rec[(size_t)p * mesh::kInts + mesh::PT + k]);
    FS_TRY(write_file(dir + "panels/leaf_point_inds_" + it, lp.data(), sizeof(int32_t) * lp.size()));
    // the whole tree (AMRStructure_restart.cpp:55-98)
    std::vector<char> buf(sizeof(int64_t) + sizeof(PanelDump) * (size_t)n_panels);
    const int64_t np = n_panels;
    std::memcpy(buf.data(), &np, sizeof np);
    for (int64_t i = 0; i < n_panels; ++i) {
        const int32_t* r = rec + (size_t)i * mesh::kInts;
        PanelDump d;
        std::memset(&d, 0, sizeof d);
        d.panel_ind = r[mesh::IND];
        d.level = r[mesh::LEVEL];
        d.parent_ind = r[mesh::PARENT];
        d.which_child = r[mesh::WHICH];
        for (int k = 0; k < 9; ++k) d.point_inds[k] = r[mesh::PT + k];
        d.left = r[mesh::LEFT];
        d.top = r[mesh::TOP];
        d.right = r[mesh::RIGHT];
        d.bottom = r[mesh::BOTTOM];
        d.child_inds_start = r[mesh::CHILD];
        d.left_bdry = (r[mesh::FLAGS] & mesh::LEFT_BDRY) ? 1 : 0;
        d.right_bdry = (r[mesh::FLAGS] & mesh::RIGHT_BDRY) ? 1 : 0;
        d.refined_xp = (r[mesh::FLAGS] & mesh::REFINED_XP) ? 1 : 0;
        d.refined_p = (r[mesh::FLAGS] & mesh::REFINED_P) ? 1 : 0;
        d.needs_refinement = (r[mesh::FLAGS] & mesh::NEEDS_REF) ? 1 : 0;
        std::memcpy(buf.data() + sizeof(int64_t) + sizeof(PanelDump) * (size_t)i, &d, sizeof d);
    }
    return write_file(dir + "panels/full_tree_" + it, buf.data(), buf.size());
}

int read_species(const char* sim_dir, const char* name, int iter, std::vector<double>& xs,
                 std::vector<double>& ps, std::vector<double>& fs, std::vector<int32_t>& rec) {
    if (!name || !*name || iter < 0) return fail(FS_E_INVALID, "read_species_output: bad argument");
    const std::string dir = species_dir(sim_dir, name), it = std::to_string(iter);
    std::vector<char> b;
    std::vector<double>* dst[3] = {&xs, &ps, &fs};
    const char* sub[3] = {"xs", "ps", "fs"};
    for (int a = 0; a < 3; ++a) {
        FS_TRY(read_file(dir + sub[a] + "/" + sub[a] + "_" + it, b));
        if (b.size() % sizeof(double))
            return fail(FS_E_INVALID, "%s_%s: size is not a whole number of doubles", sub[a],
                        it.c_str());
        dst[a]->resize(b.size() / sizeof(double));
        if (!b.empty()) std::memcpy(dst[a]->data(), b.data(), b.size());
    }
    if (xs.size() != ps.size() || ps.size() != fs.size())
        return fail(FS_E_INVALID, "mismatched xs/ps/fs sizes on restart load: %zu / %zu / %zu",
                    xs.size(), ps.size(), fs.size());
    FS_TRY(read_file(dir + "panels/full_tree_" + it, b));
    int64_t np = 0;
    if (b.size() < sizeof np) return fail(FS_E_INVALID, "full_tree_%s: no header", it.c_str());
    std::memcpy(&np, b.data(), sizeof np);
    if (np <= 0 || b.size() != sizeof np + sizeof(PanelDump) * (size_t)np)
        return fail(FS_E_INVALID, "full_tree_%s: bad header or size", it.c_str());
    rec.assign((size_t)np * mesh::kInts, 0);
    for (int64_t i = 0; i < np; ++i) {
        PanelDump d;
        std::memcpy(&d, b.data() + sizeof np + sizeof(PanelDump) * (size_t)i, sizeof d);
        int32_t* r = &rec[(size_t)i * mesh::kInts];
        r[mesh::IND] = d.panel_ind;
        for (int k = 0; k < 9; ++k) {
            if (d.point_inds[k] < 0 || d.point_inds[k] >= (int64_t)xs.size())
                return fail(FS_E_INVALID, "full_tree_%s: panel %lld names point %d of %zu",
                            it.c_str(), (long long)i, d.point_inds[k], xs.size());
            r[mesh::PT + k] = d.point_inds[k];
        }
        r[mesh::LEVEL] = d.level;
        r[mesh::PARENT] = d.parent_ind;
        r[mesh::WHICH] = d.which_child;
        r[mesh::LEFT] = d.left;
        r[mesh::TOP] = d.top;
        r[mesh::RIGHT] = d.right;
        r[mesh::BOTTOM] = d.bottom;
        r[mesh::FLAGS] = (d.left_bdry ? mesh::LEFT_BDRY : 0) | (d.right_bdry ? mesh::RIGHT_BDRY : 0) |
                         (d.needs_refinement ? mesh::NEEDS_REF : 0) |
                         (d.refined_xp ? mesh::REFINED_XP : 0) | (d.refined_p ? mesh::REFINED_P : 0);
        r[mesh::CHILD] = d.child_inds_start;
        const int nch = d.refined_p ? 2 : d.refined_xp ? 4 : 0;
        if (nch && (d.child_inds_start <= i || d.child_inds_start + nch > np))
            return fail(FS_E_INVALID, "full_tree_%s: panel %lld has children out of range",
                        it.c_str(), (long long)i);
    }
    return FS_OK;
}

int write_reduced(const char* sim_dir, int iter, int64_t u, const double* rx, const double* re) {
    std::string d = sim_dir ? sim_dir : "";
    if (!d.empty() && d.back() != '/') d += '/';
    d += "simulation_output/";
    const std::string it = std::to_string(iter);
    FS_TRY(make_dirs(d + "reduce_xs"));
    FS_TRY(make_dirs(d + "reduced_es"));
    FS_TRY(write_file(d + "reduce_xs/xs_" + it, rx, sizeof(double) * (size_t)u));
    return write_file(d + "reduced_es/es_" + it, re, sizeof(double) * (size_t)u);
}

}  // namespace io
}  // namespace fs

using namespace fs;

extern "C" {

int fs_write_species_output(const char* sim_dir, const char* species_name, int iter, int64_t n,
                            const double* xs, const double* ps, const double* fs,
                            const double* q_ws, const double* es, int64_t n_panels,
                            const int32_t* panels) {
    return io::write_species(sim_dir, species_name, iter, n, xs, ps, fs, q_ws, es, n_panels,
                             panels);
}

int fs_read_species_output(const char* sim_dir, const char* species_name, int iter,
                           int64_t* n, double* xs, double* ps, double* fs, int64_t* n_panels,
                           int32_t* panels) {
    std::vector<double> x, p, f;
    std::vector<int32_t> rec;
    FS_TRY(io::read_species(sim_dir, species_name, iter, x, p, f, rec));
    if (n) *n = (int64_t)x.size();
    if (n_panels) *n_panels = (int64_t)(rec.size() / mesh::kInts);
    if (xs) std::copy(x.begin(), x.end(), xs);
    if (ps) std::copy(p.begin(), p.end(), ps);
    if (fs) std::copy(f.begin(), f.end(), fs);
    if (panels) std::copy(rec.begin(), rec.end(), panels);
    return FS_OK;
}

}  // extern "C"
