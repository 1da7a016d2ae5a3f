// TetGen .node / .ele loading for the drop-in (the loader of reference main.cpp:14-168).
// Files are read whole and parsed in place (strtol / strtof per field) rather than through a
// stringstream per line; empty lines and lines starting with '#' are skipped.
//
// readNodeFile: first line "<#points> <dim> <#attributes> <#markers>", then one
// "<index> <x> <y> <z> ..." line per point; start_from_0 is set when the first index is 0
// (main.cpp:38-39).  loadTetgenModel reads the rest-shape nodes, optional current-shape
// nodes (empty name or unreadable file -> rest shape, main.cpp:67-72), and the .ele tets
// ("<#tets> <4> <#attributes>", then "<index> <v0> <v1> <v2> <v3> ..."), shifts 1-based ids
// to 0-based (main.cpp:112-114), pins every vertex whose coordinate on `pinAxis` is within
// `pinFraction` of the extent from the minimum (main.cpp:146-164 pins y <= min + 5%, tested
// on the current positions as the reference does), and calls TetMesh::initialize.
#pragma once

#include <cstdlib>
#include <fstream>
#include <limits>
#include <string>
#include <vector>

#include "TetMesh.h"

namespace VBD {

namespace tetgen_detail {

// Whole-file read; the std::string keeps a terminating NUL, so strtol/strtof never run past it.
inline bool slurp(const std::string& path, std::string& buf) {
    std::ifstream in(path.c_str(), std::ios::binary);
    if (!in) return false;
    in.seekg(0, std::ios::end);
    const std::streamoff n = in.tellg();
    if (n < 0) return false;
    buf.resize((size_t)n);
    in.seekg(0, std::ios::beg);
    return n == 0 || (bool)in.read(&buf[0], n);
}

// Lines of a buffer, skipping empty lines and lines that start with '#'.
struct Lines {
    const char* p;
    const char* end;
    bool next(const char*& b, const char*& e) {
        while (p < end) {
            b = p;
            while (p < end && *p != '\n') ++p;
            e = p;
            if (p < end) ++p;
            const char* q = b;
            while (q < e && (*q == ' ' || *q == '\t' || *q == '\r')) ++q;
            if (q < e && *b != '#') return true;
        }
        return false;
    }
};

// Integer / float fields of one line [b, e): fail when the line has fewer fields.
inline bool field(const char*& b, const char* e, long& v) {
    char* q = nullptr;
    v = std::strtol(b, &q, 10);
    if (q == b || q > e) return false;
    b = q;
    return true;
}
inline bool field(const char*& b, const char* e, float& v) {
    char* q = nullptr;
    v = std::strtof(b, &q);
    if (q == b || q > e) return false;
    b = q;
    return true;
}

}  // namespace tetgen_detail

inline std::vector<Vec3> readNodeFile(const std::string& nodeFilename, bool& start_from_0) {
    std::vector<Vec3> positions;
    std::string buf;
    if (!tetgen_detail::slurp(nodeFilename, buf)) return positions;
    tetgen_detail::Lines L{buf.data(), buf.data() + buf.size()};
    const char *b, *e;
    long n = 0;
    if (!L.next(b, e) || !tetgen_detail::field(b, e, n) || n < 0) return positions;
    positions.reserve((size_t)n);
    while ((long)positions.size() < n && L.next(b, e)) {
        long idx;
        float x, y, z;
        if (!tetgen_detail::field(b, e, idx) || !tetgen_detail::field(b, e, x) || !tetgen_detail::field(b, e, y) ||
            !tetgen_detail::field(b, e, z))
            break;
        if (positions.empty() && idx == 0) start_from_0 = true;
        positions.emplace_back(x, y, z);
    }
    if ((long)positions.size() != n) positions.clear();
    return positions;
}

inline bool readEleFile(const std::string& eleFilename, bool start_from_0, std::vector<TetVIdsArr>& tets) {
    std::string buf;
    if (!tetgen_detail::slurp(eleFilename, buf)) return false;
    tetgen_detail::Lines L{buf.data(), buf.data() + buf.size()};
    const char *b, *e;
    long n = 0;
    if (!L.next(b, e) || !tetgen_detail::field(b, e, n) || n < 0) return false;
    tets.clear();
    tets.reserve((size_t)n);
    while ((long)tets.size() < n && L.next(b, e)) {
        long idx, v[4];
        if (!tetgen_detail::field(b, e, idx)) return false;
        for (long& x : v)
            if (!tetgen_detail::field(b, e, x)) return false;
        TetVIdsArr t;
        for (int k = 0; k < 4; ++k) t[k] = (IdType)(start_from_0 ? v[k] : v[k] - 1);
        tets.push_back(t);
    }
    return (long)tets.size() == n;
}

inline bool loadTetgenModel(const std::string& restNodeFilename, const std::string& currentNodeFilename,
                            const std::string& eleFilename, TetMesh& mesh, Real density, Real mu, Real lambda,
                            int pinAxis = 1, Real pinFraction = 0.05f) {
    bool start_from_0 = false;
    std::vector<Vec3> rest = readNodeFile(restNodeFilename, start_from_0);
    if (rest.empty()) return false;
    bool dummy = false;
    std::vector<Vec3> cur = currentNodeFilename.empty() ? std::vector<Vec3>() : readNodeFile(currentNodeFilename, dummy);
    if (cur.empty()) cur = rest;
    if (cur.size() != rest.size()) return false;
    std::vector<TetVIdsArr> tets;
    if (!readEleFile(eleFilename, start_from_0, tets)) return false;
    for (const auto& t : tets)
        for (IdType v : t)
            if (v < 0 || v >= (IdType)rest.size()) return false;
    std::vector<IdType> fixedVerts;
    if (pinAxis >= 0 && pinAxis < 3) {
        Real lo = std::numeric_limits<Real>::max(), hi = std::numeric_limits<Real>::lowest();
        for (const auto& p : cur) {
            lo = std::min(lo, p[pinAxis]);
            hi = std::max(hi, p[pinAxis]);
        }
        const Real thr = lo + (hi - lo) * pinFraction;
        for (size_t i = 0; i < cur.size(); ++i)
            if (cur[i][pinAxis] <= thr) fixedVerts.push_back((IdType)i);
    }
    mesh.initialize(rest, cur, density, tets, mu, lambda, fixedVerts);
    return true;
}

}  // namespace VBD
