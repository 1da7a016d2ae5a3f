// TetGen .node / .ele loading for the drop-in (the loader of reference main.cpp:14-168).
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

#include <fstream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "TetMesh.h"

namespace VBD {

inline std::vector<Vec3> readNodeFile(const std::string& nodeFilename, bool& start_from_0) {
    std::vector<Vec3> positions;
    std::ifstream in(nodeFilename.c_str());
    if (!in) return positions;
    std::string line;
    size_t n = 0;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        ss >> n;
        break;
    }
    positions.reserve(n);
    while (positions.size() < n && std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        long idx;
        Real x, y, z;
        if (!(ss >> idx >> x >> y >> z)) break;
        if (positions.empty() && idx == 0) start_from_0 = true;
        positions.emplace_back(x, y, z);
    }
    if (positions.size() != n) positions.clear();
    return positions;
}

inline bool readEleFile(const std::string& eleFilename, bool start_from_0, std::vector<TetVIdsArr>& tets) {
    std::ifstream in(eleFilename.c_str());
    if (!in) return false;
    std::string line;
    size_t n = 0;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        ss >> n;
        break;
    }
    tets.clear();
    tets.reserve(n);
    while (tets.size() < n && std::getline(in, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ss(line);
        long idx;
        TetVIdsArr t;
        if (!(ss >> idx >> t[0] >> t[1] >> t[2] >> t[3])) return false;
        if (!start_from_0)
            for (auto& v : t) v -= 1;
        tets.push_back(t);
    }
    return tets.size() == n;
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
