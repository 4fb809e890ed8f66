/*
 * mesh_io.cpp -- Wavefront OBJ import without assimp.  The reference imports
 * meshes through assimp with triangulation, smooth normals, flipped UVs and
 * tangent space (src/delegate/meshimporter.cpp:41-121); this reader does the
 * same for OBJ:
 *
 *   - "v", "vt", "vn" and "f" records (1-based or negative indices, forms
 *     v, v/vt, v//vn, v/vt/vn); other records are ignored; all groups and
 *     objects merge into one mesh;
 *   - polygons are fanned into triangles (v0, vk, vk+1);
 *   - each distinct (v, vt, vn) corner becomes one output vertex;
 *   - without "vn", normals are the angle-independent smooth normals: the
 *     sum over incident triangles of the unnormalised face normal (area
 *     weighted), normalised, shared by every corner at a position;
 *   - texture v is flipped (v -> 1 - v);
 *   - with texture coordinates, per-vertex tangents are the sum of the
 *     triangles' dP/du (from the UV parametrisation), made orthogonal to the
 *     normal (Gram-Schmidt) and normalised; degenerate cases fall back to
 *     any unit vector orthogonal to the normal.
 */
#include "../../include/clhip.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

namespace
{
thread_local std::string g_mesh_error;

int mesh_fail(const std::string& m)
{
    g_mesh_error = m;
    return CLHIP_E_INVALID;
}

struct V3
{
    double x = 0, y = 0, z = 0;
};
V3     sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3     add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3     scl(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3     cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3     normalize(V3 a)
{
    const double l = std::sqrt(dot(a, a));
    return l > 0 ? scl(a, 1.0 / l) : V3{0, 0, 1};
}

struct ObjMesh
{
    std::vector<float>    pos, nrm, tan, uv;
    std::vector<uint32_t> idx;
    bool                  has_uv = false;
};

/* index of record i (1-based or negative) into a list of n */
bool resolve(long i, size_t n, size_t& out)
{
    if (i > 0 && (size_t)i <= n)
        out = (size_t)i - 1;
    else if (i < 0 && (size_t)(-i) <= n)
        out = n - (size_t)(-i);
    else
        return false;
    return true;
}

int parse_obj(const char* path, ObjMesh& m)
{
    FILE* f = std::fopen(path, "rb");
    if (!f)
        return mesh_fail(std::string("cannot open ") + path);
    std::vector<V3>                                         P, N;
    std::vector<std::pair<double, double>>                  T;
    std::map<std::tuple<size_t, size_t, size_t>, uint32_t>  corner; /* (v, vt+1, vn+1) -> vertex */
    std::vector<size_t>                                     vpos;   /* output vertex -> position index */
    std::vector<long>                                       vuv, vn;
    char                                                    buf[4096];
    int                                                     lineno = 0;
    bool                                                    any_vn = false, any_vt = false;
    while (std::fgets(buf, sizeof buf, f))
    {
        ++lineno;
        char* s = buf;
        while (*s == ' ' || *s == '\t')
            ++s;
        if (s[0] == 'v' && (s[1] == ' ' || s[1] == '\t'))
        {
            V3 p;
            if (std::sscanf(s + 2, "%lf %lf %lf", &p.x, &p.y, &p.z) != 3)
                break;
            P.push_back(p);
        }
        else if (s[0] == 'v' && s[1] == 'n')
        {
            V3 n;
            if (std::sscanf(s + 2, "%lf %lf %lf", &n.x, &n.y, &n.z) != 3)
                break;
            N.push_back(n);
        }
        else if (s[0] == 'v' && s[1] == 't')
        {
            double u = 0, v = 0;
            if (std::sscanf(s + 2, "%lf %lf", &u, &v) < 1)
                break;
            T.push_back({u, v});
        }
        else if (s[0] == 'f' && (s[1] == ' ' || s[1] == '\t'))
        {
            std::vector<uint32_t> face;
            char*                 tok = std::strtok(s + 1, " \t\r\n");
            while (tok)
            {
                long  iv = 0, it = 0, in = 0;
                char* e  = nullptr;
                iv       = std::strtol(tok, &e, 10);
                if (*e == '/')
                {
                    ++e;
                    if (*e != '/')
                        it = std::strtol(e, &e, 10);
                    if (*e == '/')
                        in = std::strtol(e + 1, &e, 10);
                }
                size_t pi = 0, ti = 0, ni = 0;
                if (!resolve(iv, P.size(), pi) || (it && !resolve(it, T.size(), ti)) || (in && !resolve(in, N.size(), ni)))
                {
                    std::fclose(f);
                    return mesh_fail(std::string(path) + ":" + std::to_string(lineno) + ": index out of range");
                }
                any_vt |= it != 0;
                any_vn |= in != 0;
                const auto key = std::make_tuple(pi, it ? ti + 1 : 0, in ? ni + 1 : 0);
                auto       fnd = corner.find(key);
                uint32_t   vid;
                if (fnd == corner.end())
                {
                    vid = (uint32_t)vpos.size();
                    corner.emplace(key, vid);
                    vpos.push_back(pi);
                    vuv.push_back(it ? (long)ti : -1);
                    vn.push_back(in ? (long)ni : -1);
                }
                else
                    vid = fnd->second;
                face.push_back(vid);
                tok = std::strtok(nullptr, " \t\r\n");
            }
            if (face.size() < 3)
            {
                std::fclose(f);
                return mesh_fail(std::string(path) + ":" + std::to_string(lineno) + ": face with fewer than 3 corners");
            }
            for (size_t k = 1; k + 1 < face.size(); ++k)
            {
                m.idx.push_back(face[0]);
                m.idx.push_back(face[k]);
                m.idx.push_back(face[k + 1]);
            }
        }
    }
    const bool bad = !std::feof(f);
    std::fclose(f);
    if (bad)
        return mesh_fail(std::string(path) + ":" + std::to_string(lineno) + ": malformed record");
    if (m.idx.empty())
        return mesh_fail(std::string(path) + ": no faces");
    const size_t nv = vpos.size();
    m.has_uv        = any_vt;
    m.pos.resize(3 * nv);
    m.nrm.resize(3 * nv);
    for (size_t i = 0; i < nv; ++i)
    {
        const V3 p   = P[vpos[i]];
        m.pos[3 * i] = (float)p.x;
        m.pos[3 * i + 1] = (float)p.y;
        m.pos[3 * i + 2] = (float)p.z;
    }
    /* normals: given, else area-weighted smooth normals per position */
    std::vector<V3> pn(P.size());
    if (!any_vn)
        for (size_t t = 0; t + 2 < m.idx.size(); t += 3)
        {
            const size_t a = vpos[m.idx[t]], b = vpos[m.idx[t + 1]], c = vpos[m.idx[t + 2]];
            const V3     fn = cross(sub(P[b], P[a]), sub(P[c], P[a]));
            pn[a]           = add(pn[a], fn);
            pn[b]           = add(pn[b], fn);
            pn[c]           = add(pn[c], fn);
        }
    for (size_t i = 0; i < nv; ++i)
    {
        const V3 n = normalize(any_vn && vn[i] >= 0 ? N[(size_t)vn[i]] : pn[vpos[i]]);
        m.nrm[3 * i]     = (float)n.x;
        m.nrm[3 * i + 1] = (float)n.y;
        m.nrm[3 * i + 2] = (float)n.z;
    }
    if (m.has_uv)
    {
        m.uv.resize(2 * nv);
        for (size_t i = 0; i < nv; ++i)
        {
            const auto t = vuv[i] >= 0 ? T[(size_t)vuv[i]] : std::make_pair(0.0, 0.0);
            m.uv[2 * i]     = (float)t.first;
            m.uv[2 * i + 1] = (float)(1.0 - t.second);
        }
        std::vector<V3> tg(nv);
        for (size_t t = 0; t + 2 < m.idx.size(); t += 3)
        {
            const uint32_t i0 = m.idx[t], i1 = m.idx[t + 1], i2 = m.idx[t + 2];
            const V3       e1 = sub(P[vpos[i1]], P[vpos[i0]]), e2 = sub(P[vpos[i2]], P[vpos[i0]]);
            const double   du1 = m.uv[2 * i1] - m.uv[2 * i0], dv1 = m.uv[2 * i1 + 1] - m.uv[2 * i0 + 1];
            const double   du2 = m.uv[2 * i2] - m.uv[2 * i0], dv2 = m.uv[2 * i2 + 1] - m.uv[2 * i0 + 1];
            const double   det = du1 * dv2 - du2 * dv1;
            if (std::fabs(det) < 1e-20)
                continue;
            const V3 dpdu = scl(sub(scl(e1, dv2), scl(e2, dv1)), 1.0 / det);
            for (uint32_t i : {i0, i1, i2})
                tg[i] = add(tg[i], dpdu);
        }
        m.tan.resize(3 * nv);
        for (size_t i = 0; i < nv; ++i)
        {
            const V3 n = {m.nrm[3 * i], m.nrm[3 * i + 1], m.nrm[3 * i + 2]};
            V3       t = sub(tg[i], scl(n, dot(tg[i], n)));
            if (dot(t, t) < 1e-24)
                t = std::fabs(n.x) > 0.9 ? cross(n, V3{0, 1, 0}) : cross(n, V3{1, 0, 0});
            t                = normalize(t);
            m.tan[3 * i]     = (float)t.x;
            m.tan[3 * i + 1] = (float)t.y;
            m.tan[3 * i + 2] = (float)t.z;
        }
    }
    return CLHIP_OK;
}
} // namespace

extern "C" {

const char* clhip_mesh_last_error(void) { return g_mesh_error.c_str(); }

int clhip_mesh_load_obj(const char* path, float* vertices, float* normals, float* tangents, float* uvs,
                        uint32_t* indices, uint32_t* num_vertices, uint32_t* num_triangles, uint32_t* has_uvs)
{
    if (!path || !num_vertices || !num_triangles)
        return mesh_fail("clhip_mesh_load_obj: bad arguments");
    ObjMesh m;
    int     rc = parse_obj(path, m);
    if (rc)
        return rc;
    const uint32_t nv = (uint32_t)(m.pos.size() / 3), nt = (uint32_t)(m.idx.size() / 3);
    const bool     query = !vertices;
    if (!query && (*num_vertices < nv || *num_triangles < nt))
        return mesh_fail("clhip_mesh_load_obj: output arrays smaller than the counts of the size query");
    *num_vertices  = nv;
    *num_triangles = nt;
    if (has_uvs)
        *has_uvs = m.has_uv ? 1u : 0u;
    if (query)
        return CLHIP_OK;
    std::memcpy(vertices, m.pos.data(), m.pos.size() * sizeof(float));
    if (indices)
        std::memcpy(indices, m.idx.data(), m.idx.size() * sizeof(uint32_t));
    if (normals)
        std::memcpy(normals, m.nrm.data(), m.nrm.size() * sizeof(float));
    if (m.has_uv && tangents)
        std::memcpy(tangents, m.tan.data(), m.tan.size() * sizeof(float));
    if (m.has_uv && uvs)
        std::memcpy(uvs, m.uv.data(), m.uv.size() * sizeof(float));
    return CLHIP_OK;
}

} // extern "C"
