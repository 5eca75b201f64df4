// ref_tinyobj.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Loads an OBJ file with the tinyobjloader that CRender vendors (external/tinyobj/tinobj.h,
// v2.0.0), configured as asset_loader.cpp:187-194 configures it (ObjReaderConfig with
// triangulate = true), and prints what the reference's loader reads from it: the vertex
// count, then one line per triangle of every shape in order,
//   "f <v0> <v1> <v2> <t0> <t1> <t2> <material>"
// (vertex and texcoord indices as tinyobj returns them, -1 for a missing texcoord).
// Built by oracle/Makefile into oracle/_ref/ from the reference's own header; the tests
// compare crt_obj_load's triangulation and primID order with it.
#define TINYOBJLOADER_IMPLEMENTATION
#include "tinobj.h"

#include <cstdio>

int main(int argc, char **argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s file.obj\n", argv[0]);
        return 2;
    }
    tinyobj::ObjReaderConfig cfg;
    cfg.triangulate = true;
    tinyobj::ObjReader reader;
    if (!reader.ParseFromFile(argv[1], cfg) && !reader.Error().empty()) {
        std::fprintf(stderr, "%s\n", reader.Error().c_str());
        return 1;
    }
    const auto &attrib = reader.GetAttrib();
    std::printf("v %zu\n", attrib.vertices.size() / 3);
    for (const auto &shape : reader.GetShapes()) {
        const auto &m = shape.mesh;
        size_t      base = 0;
        for (size_t f = 0; f < m.num_face_vertices.size(); ++f) {
            const unsigned nv = m.num_face_vertices[f];
            if (nv == 3) {
                const auto &a = m.indices[base], &b = m.indices[base + 1], &c = m.indices[base + 2];
                std::printf("f %d %d %d %d %d %d %d\n", a.vertex_index, b.vertex_index, c.vertex_index,
                            a.texcoord_index, b.texcoord_index, c.texcoord_index, m.material_ids[f]);
            } else {
                std::printf("n %u\n", nv);   // a face that was not triangulated
            }
            base += nv;
        }
    }
    return 0;
}
