// BookShelf placement files: reader (.aux/.scl/.nodes/.nets/.pl/.wts) and
// .pl writer, producing the SoA + CSR database the engine takes.
//
// Semantics follow the reference's BookshelfReader::read and
// BookshelfWriter::writePl (src/io/bookshelf_reader.cpp:130-463,
// src/io/bookshelf_writer.cpp:16-37) and PlaceDB's partition / CSR / region
// rules (src/db/place_db.cpp:31-125):
//  * .pl holds lower-left corners, the database centres (one conversion each
//    way, in float on reading and in double on writing);
//  * nodes are stably reordered to [movable | fixed]; a node is fixed when
//    .nodes says terminal / terminal_NI or .pl says /FIXED;
//  * a movable node taller than 1.5 row heights is a macro; the row height
//    is the most frequent row height (ties: the smallest);
//  * net->pin and node->pin CSR lists keep pin order inside each list.
#include <algorithm>
#include <charconv>
#include <cstdio>
#include <fstream>
#include <map>
#include <memory>
#include <cstring>
#include <string>
#include <string_view>
#include <unordered_map>
#include <vector>

#include "abi.h"

namespace sphip {
namespace {

// A whole text file, handed out line by line with surrounding blanks cut.
class TextLines {
public:
    explicit TextLines(const std::string& path) {
        std::ifstream in(path, std::ios::binary);
        if (!in) throw InvalidArg("cannot open " + path);
        text_.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
    }
    // Next non-empty line that is not a comment or a "UCLA ..." header,
    // split into whitespace-separated fields.  False at the end.
    bool next(std::vector<std::string_view>& fields) {
        while (pos_ < text_.size()) {
            size_t end = text_.find('\n', pos_);
            if (end == std::string::npos) end = text_.size();
            std::string_view line(text_.data() + pos_, end - pos_);
            pos_ = end + 1;
            split(line, fields);
            if (fields.empty() || fields[0][0] == '#' || fields[0] == "UCLA") continue;
            return true;
        }
        return false;
    }

private:
    static bool blank(char c) { return c == ' ' || c == '\t' || c == '\r'; }
    static void split(std::string_view line, std::vector<std::string_view>& out) {
        out.clear();
        size_t i = 0;
        while (i < line.size()) {
            while (i < line.size() && blank(line[i])) ++i;
            const size_t b = i;
            while (i < line.size() && !blank(line[i])) ++i;
            if (i > b) out.push_back(line.substr(b, i - b));
        }
    }
    std::string text_;
    size_t pos_ = 0;
};

float parseFloat(std::string_view s) {
    float v = 0.f;
    const auto r = std::from_chars(s.data(), s.data() + s.size(), v);
    if (r.ec != std::errc()) throw InvalidArg("bad number '" + std::string(s) + "'");
    return v;
}

int parseInt(std::string_view s) {
    int v = 0;
    const auto r = std::from_chars(s.data(), s.data() + s.size(), v);
    if (r.ec == std::errc()) return v;
    return static_cast<int>(static_cast<double>(parseFloat(s)));   // "12.000000"
}

// The value after ':' in "Key : value" (also "Key:value" and "Key: value"),
// searching from field `from`.
bool valueAfterColon(const std::vector<std::string_view>& f, size_t from, std::string_view* v) {
    for (size_t i = from; i < f.size(); ++i) {
        if (f[i] == ":") {
            if (i + 1 < f.size()) {
                *v = f[i + 1];
                return true;
            }
            return false;
        }
        const size_t c = f[i].find(':');
        if (c != std::string_view::npos && c + 1 < f[i].size()) {
            *v = f[i].substr(c + 1);
            return true;
        }
    }
    return false;
}

const char* const kOrients[8] = {"N", "E", "S", "W", "FN", "FE", "FS", "FW"};

uint8_t orientCode(std::string_view s) {
    for (uint8_t k = 0; k < 8; ++k)
        if (s == kOrients[k]) return k;
    return 0;
}

std::string extensionOf(std::string_view name) {
    const size_t dot = name.rfind('.');
    return dot == std::string_view::npos ? std::string() : std::string(name.substr(dot));
}

std::string joinDir(const std::string& dir, std::string_view name) {
    return dir.empty() ? std::string(name) : dir + "/" + std::string(name);
}

}  // namespace

// The database of one BookShelf design (node order [movable | fixed]).
struct Bookshelf {
    int numMovable = 0, numNodes = 0;
    std::vector<std::string> names, netNames;
    std::vector<float> x, y, w, h;
    std::vector<uint8_t> flags, orient;
    std::vector<float> offX, offY, netWeight;
    std::vector<int32_t> pin2node, pin2net, netStart, netPins, nodeStart, nodePins;
    std::vector<float> rowLy, rowHeight, rowLx, rowHx, rowStep;
    std::vector<int32_t> rowSites;
    float coreLx = 0, coreLy = 0, coreHx = 0, coreHy = 0, rowH = 0;

    void read(const std::string& auxPath);
    void readPlacement(const std::string& plPath);
    void writePl(const std::string& path, const float* cx, const float* cy,
                 const uint8_t* nodeFlags = nullptr, const uint8_t* nodeOrient = nullptr) const;

private:
    std::unordered_map<std::string_view, int> nodeIndex() const {
        std::unordered_map<std::string_view, int> m;
        m.reserve(names.size() * 2);
        for (int i = 0; i < static_cast<int>(names.size()); ++i) m.emplace(names[i], i);
        return m;
    }
    void readScl(const std::string& path);
    void readNodes(const std::string& path);
    void readNets(const std::string& path);
    void readWts(const std::string& path);
    void regions();
};

void Bookshelf::read(const std::string& auxPath) {
    std::string nodesP, netsP, wtsP, plP, sclP;
    {
        TextLines aux(auxPath);
        const size_t slash = auxPath.rfind('/');
        const std::string dir = slash == std::string::npos ? "." : auxPath.substr(0, slash);
        std::vector<std::string_view> f;
        while (aux.next(f))
            for (std::string_view t : f) {
                const std::string ext = extensionOf(t);
                if (t == ":" || ext.empty()) continue;
                if (ext == ".nodes") nodesP = joinDir(dir, t);
                else if (ext == ".nets") netsP = joinDir(dir, t);
                else if (ext == ".wts") wtsP = joinDir(dir, t);
                else if (ext == ".pl") plP = joinDir(dir, t);
                else if (ext == ".scl") sclP = joinDir(dir, t);
            }
        if (nodesP.empty() || netsP.empty() || plP.empty() || sclP.empty())
            throw InvalidArg(".aux lacks a .nodes, .nets, .pl or .scl entry: " + auxPath);
    }
    readScl(sclP);
    regions();
    readNodes(nodesP);
    readNets(netsP);
    readPlacement(plP);
    readWts(wtsP);
    for (int i = 0; i < numNodes; ++i)
        if (!(flags[i] & F_FIXED) && h[i] > rowH * 1.5f) flags[i] |= F_MACRO;

    // stable [movable | fixed] order
    std::vector<int> order;
    order.reserve(numNodes);
    for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < numNodes; ++i)
            if (((flags[i] & F_FIXED) != 0) == (pass == 1)) order.push_back(i);
    numMovable = 0;
    for (int i = 0; i < numNodes; ++i)
        if (!(flags[i] & F_FIXED)) ++numMovable;
    std::vector<int> newIndex(numNodes);
    for (int k = 0; k < numNodes; ++k) newIndex[order[k]] = k;
    auto reorder = [&](auto& v) {
        std::remove_reference_t<decltype(v)> t(v.size());
        for (int k = 0; k < numNodes; ++k) t[k] = std::move(v[order[k]]);
        v.swap(t);
    };
    reorder(names);
    reorder(x);
    reorder(y);
    reorder(w);
    reorder(h);
    reorder(flags);
    reorder(orient);
    for (int32_t& n : pin2node) n = newIndex[n];

    // CSR lists, pins in index order inside each list
    const int P = static_cast<int>(pin2node.size()), E = static_cast<int>(netNames.size());
    auto csr = [P](const std::vector<int32_t>& owner, int owners, std::vector<int32_t>& start,
                   std::vector<int32_t>& list) {
        start.assign(owners + 1, 0);
        for (int p = 0; p < P; ++p) ++start[owner[p] + 1];
        for (int k = 0; k < owners; ++k) start[k + 1] += start[k];
        list.resize(P);
        std::vector<int32_t> cur(start.begin(), start.end() - 1);
        for (int p = 0; p < P; ++p) list[cur[owner[p]]++] = p;
    };
    csr(pin2net, E, netStart, netPins);
    csr(pin2node, numNodes, nodeStart, nodePins);
}

void Bookshelf::readScl(const std::string& path) {
    TextLines in(path);
    std::vector<std::string_view> f;
    bool open = false;
    float ly = 0, height = 0, lx = 0, siteW = 1, siteSp = 0;
    int sites = 0;
    auto close = [&] {
        const float step = siteSp > 0.f ? siteSp : siteW;
        rowLy.push_back(ly);
        rowHeight.push_back(height);
        rowLx.push_back(lx);
        rowStep.push_back(step);
        rowSites.push_back(sites);
        rowHx.push_back(lx + step * static_cast<float>(sites));
        open = false;
    };
    while (in.next(f)) {
        const std::string_view key = f[0];
        std::string_view v;
        if (key == "CoreRow") {
            open = true;
            ly = height = lx = 0;
            siteW = 1;
            siteSp = 0;
            sites = 0;
        } else if (key == "End") {
            if (open) close();
        } else if (!open) {
            continue;
        } else if (key == "Coordinate") {
            if (valueAfterColon(f, 0, &v)) ly = parseFloat(v);
        } else if (key == "Height") {
            if (valueAfterColon(f, 0, &v)) height = parseFloat(v);
        } else if (key == "Sitewidth") {
            if (valueAfterColon(f, 0, &v)) siteW = parseFloat(v);
        } else if (key == "Sitespacing") {
            if (valueAfterColon(f, 0, &v)) siteSp = parseFloat(v);
        } else if (key == "SubrowOrigin") {
            // "SubrowOrigin : lx  NumSites : n" on one line
            if (valueAfterColon(f, 0, &v)) lx = parseFloat(v);
            for (size_t i = 1; i < f.size(); ++i)
                if (f[i].substr(0, 8) == "NumSites") {
                    if (valueAfterColon(f, i, &v)) sites = parseInt(v);
                    break;
                }
        }
    }
    if (open) close();
    if (rowLy.empty()) throw InvalidArg("no CoreRow in " + path);
}

// Core = bounding box of the rows; row height = most frequent height.
void Bookshelf::regions() {
    std::map<float, int> count;
    for (float hh : rowHeight) ++count[hh];
    int best = -1;
    for (const auto& [hh, c] : count)
        if (c > best) {
            best = c;
            rowH = hh;
        }
    coreLx = rowLx[0];
    coreLy = rowLy[0];
    coreHx = rowHx[0];
    coreHy = rowLy[0] + rowHeight[0];
    for (size_t r = 1; r < rowLy.size(); ++r) {
        coreLx = std::min(coreLx, rowLx[r]);
        coreLy = std::min(coreLy, rowLy[r]);
        coreHx = std::max(coreHx, rowHx[r]);
        coreHy = std::max(coreHy, rowLy[r] + rowHeight[r]);
    }
}

void Bookshelf::readNodes(const std::string& path) {
    TextLines in(path);
    std::vector<std::string_view> f;
    while (in.next(f)) {
        if (f[0] == "NumNodes" || f[0] == "NumTerminals") continue;
        if (f.size() < 3) continue;
        uint8_t fl = 0;
        if (f.size() >= 4) {
            if (f[3] == "terminal") fl = F_FIXED;
            else if (f[3] == "terminal_NI") fl = F_FIXED | F_NI;
        }
        names.emplace_back(f[0]);
        w.push_back(parseFloat(f[1]));
        h.push_back(parseFloat(f[2]));
        flags.push_back(fl);
    }
    numNodes = static_cast<int>(names.size());
    x.assign(numNodes, 0.f);
    y.assign(numNodes, 0.f);
    orient.assign(numNodes, 0);
}

void Bookshelf::readNets(const std::string& path) {
    TextLines in(path);
    std::vector<std::string_view> f;
    const auto index = nodeIndex();
    int left = 0;
    while (in.next(f)) {
        if (f[0] == "NumNets" || f[0] == "NumPins") continue;
        if (f[0] == "NetDegree") {
            std::string_view v;
            if (!valueAfterColon(f, 0, &v)) throw InvalidArg("NetDegree line without a degree");
            left = parseInt(v);
            const int k = static_cast<int>(netNames.size());
            std::string name = "n" + std::to_string(k);
            for (size_t i = 0; i + 2 < f.size(); ++i)
                if (f[i] == ":") {
                    name = std::string(f[i + 2]);
                    break;
                }
            netNames.push_back(std::move(name));
            continue;
        }
        if (netNames.empty() || left <= 0) continue;
        // "node [I|O|B] : dx dy"
        const auto it = index.find(f[0]);
        if (it == index.end()) throw InvalidArg("net pin on unknown node " + std::string(f[0]));
        float dx = 0.f, dy = 0.f;
        for (size_t i = 1; i < f.size(); ++i)
            if (f[i] == ":") {
                if (i + 2 < f.size()) {
                    dx = parseFloat(f[i + 1]);
                    dy = parseFloat(f[i + 2]);
                }
                break;
            }
        pin2node.push_back(it->second);
        pin2net.push_back(static_cast<int32_t>(netNames.size()) - 1);
        offX.push_back(dx);
        offY.push_back(dy);
        --left;
    }
}

void Bookshelf::readPlacement(const std::string& path) {
    TextLines in(path);
    std::vector<std::string_view> f;
    const auto index = nodeIndex();
    while (in.next(f)) {
        if (f.size() < 3) continue;
        const auto it = index.find(f[0]);
        if (it == index.end()) continue;
        const int i = it->second;
        x[i] = parseFloat(f[1]) + 0.5f * w[i];
        y[i] = parseFloat(f[2]) + 0.5f * h[i];
        for (size_t k = 3; k < f.size(); ++k) {
            if (f[k] == ":") {
                if (k + 1 < f.size()) orient[i] = orientCode(f[k + 1]);
            } else if (f[k].find("FIXED") != std::string_view::npos) {
                flags[i] |= F_FIXED;
            }
        }
    }
}

void Bookshelf::readWts(const std::string& path) {
    netWeight.assign(netNames.size(), 1.0f);
    if (path.empty()) return;
    std::ifstream probe(path);
    if (!probe) return;   // optional file
    TextLines in(path);
    std::unordered_map<std::string_view, int> index;
    index.reserve(netNames.size() * 2);
    for (int k = 0; k < static_cast<int>(netNames.size()); ++k) index.emplace(netNames[k], k);
    std::vector<std::string_view> f;
    while (in.next(f)) {
        if (f.size() < 2) continue;
        const auto it = index.find(f[0]);
        if (it != index.end()) netWeight[it->second] = parseFloat(f[1]);
    }
}

void Bookshelf::writePl(const std::string& path, const float* cx, const float* cy,
                        const uint8_t* nodeFlags, const uint8_t* nodeOrient) const {
    const uint8_t* fl = nodeFlags ? nodeFlags : flags.data();
    const uint8_t* orn = nodeOrient ? nodeOrient : orient.data();
    std::FILE* out = std::fopen(path.c_str(), "wb");
    if (!out) throw InvalidArg("cannot write " + path);
    std::fputs("UCLA pl 1.0\n# Generated by superPlace\n\n", out);
    for (int i = 0; i < numNodes; ++i) {
        const double llx = static_cast<double>(cx[i]) - 0.5 * w[i];
        const double lly = static_cast<double>(cy[i]) - 0.5 * h[i];
        std::fprintf(out, "%s\t%.6f\t%.6f\t: %s%s\n", names[i].c_str(), llx, lly,
                     kOrients[orn[i] < 8 ? orn[i] : 0], (fl[i] & F_FIXED) ? " /FIXED" : "");
    }
    if (std::fclose(out) != 0) throw InvalidArg("cannot write " + path);
}

}  // namespace sphip

struct sp_bookshelf {
    sphip::Bookshelf db;
};

extern "C" {

using sphip::guarded;
using sphip::InvalidArg;

int sp_bookshelf_read(const char* aux_path, sp_bookshelf** out) {
    return guarded([&] {
        if (!aux_path || !out) throw InvalidArg("null argument");
        *out = nullptr;
        auto b = std::make_unique<sp_bookshelf>();
        b->db.read(aux_path);
        *out = b.release();
    });
}

int sp_bookshelf_design(const sp_bookshelf* b, sp_hip_design* d) {
    return guarded([&] {
        if (!b || !d) throw InvalidArg("null argument");
        const sphip::Bookshelf& s = b->db;
        *d = sp_hip_design{};
        d->num_movable = s.numMovable;
        d->num_nodes = s.numNodes;
        d->num_fillers = 0;
        d->num_pins = static_cast<int32_t>(s.pin2node.size());
        d->num_nets = static_cast<int32_t>(s.netNames.size());
        d->num_rows = static_cast<int32_t>(s.rowLy.size());
        d->node_x = s.x.data();
        d->node_y = s.y.data();
        d->node_w = s.w.data();
        d->node_h = s.h.data();
        d->node_flags = s.flags.data();
        d->pin_offset_x = s.offX.data();
        d->pin_offset_y = s.offY.data();
        d->pin2node = s.pin2node.data();
        d->pin2net = s.pin2net.data();
        d->net2pin_start = s.netStart.data();
        d->flat_net2pin = s.netPins.data();
        d->node2pin_start = s.nodeStart.data();
        d->flat_node2pin = s.nodePins.data();
        d->row_ly = s.rowLy.data();
        d->row_height = s.rowHeight.data();
        d->row_lx = s.rowLx.data();
        d->row_hx = s.rowHx.data();
        d->core_lx = s.coreLx;
        d->core_ly = s.coreLy;
        d->core_hx = s.coreHx;
        d->core_hy = s.coreHy;
    });
}

int sp_bookshelf_extra(const sp_bookshelf* b, const uint8_t** orient, const float** net_weight,
                       const float** row_step, const int32_t** row_sites, float* row_height) {
    return guarded([&] {
        if (!b) throw InvalidArg("null argument");
        if (orient) *orient = b->db.orient.data();
        if (net_weight) *net_weight = b->db.netWeight.data();
        if (row_step) *row_step = b->db.rowStep.data();
        if (row_sites) *row_sites = b->db.rowSites.data();
        if (row_height) *row_height = b->db.rowH;
    });
}

int sp_bookshelf_names(const sp_bookshelf* b, int which, char* buf, int64_t cap, int64_t* need) {
    return guarded([&] {
        if (!b || (which != 0 && which != 1)) throw InvalidArg("bad argument");
        const auto& v = which == 0 ? b->db.names : b->db.netNames;
        int64_t n = 0;
        for (const auto& s : v) n += static_cast<int64_t>(s.size()) + 1;
        if (need) *need = n;
        if (!buf) return;
        if (cap < n) throw InvalidArg("name buffer too small");
        char* p = buf;
        for (const auto& s : v) {
            std::memcpy(p, s.data(), s.size());
            p += s.size();
            *p++ = '\n';
        }
    });
}

int sp_bookshelf_read_pl(sp_bookshelf* b, const char* pl_path) {
    return guarded([&] {
        if (!b || !pl_path) throw InvalidArg("null argument");
        b->db.readPlacement(pl_path);
    });
}

int sp_bookshelf_write_pl(const sp_bookshelf* b, const char* path, const float* x,
                          const float* y) {
    return guarded([&] {
        if (!b || !path || !x || !y) throw InvalidArg("null argument");
        b->db.writePl(path, x, y);
    });
}

int sp_bookshelf_write_pl_state(const sp_bookshelf* b, const char* path, const float* x,
                                const float* y, const uint8_t* node_flags,
                                const uint8_t* node_orient) {
    return guarded([&] {
        if (!b || !path || !x || !y) throw InvalidArg("null argument");
        b->db.writePl(path, x, y, node_flags, node_orient);
    });
}

void sp_bookshelf_destroy(sp_bookshelf* b) { delete b; }

}  // extern "C"
