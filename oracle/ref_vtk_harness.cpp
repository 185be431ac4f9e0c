// ref_vtk_harness.cpp -- test infrastructure only.  Drives the reference's own VTK
// XML writer (VtkXmlWriter.h with its vendored tinyxml2 and cppcodec, read in place
// from the reference checkout) through the call sequences of Points::write_vtk
// (Points.h:851-1039) and Surfaces::write_vtk (Surfaces.h:1301-1450).  The sequences
// are restated here because Points.h and Surfaces.h need Eigen3 (through
// ElementBase.h -> Body.h).  Built by build_ref.sh into
// oracle/_ref/libo3dref_vtk.so; the product never loads it.
// VtkXmlWriter.h uses std::cout and std::unique_ptr without including their
// standard headers (its includers in the reference provide them)
#include <iostream>
#include <memory>

#include "VectorHelper.h"
#include "VtkXmlWriter.h"

#include <array>
#include <cstdint>
#include <map>
#include <numeric>
#include <string>

extern "C" int ref_write_points_vtk(const char* fn, int64_t n, const float* x, const float* y, const float* z,
                                    const float* sx, const float* sy, const float* sz, const float* r,
                                    const float* u, const float* v, const float* w, double time, int inert) {
  const bool has_strengths = !inert, has_radii = !inert;
  VtkXmlWriter ptsWriter = VtkXmlWriter(fn, true);
  ptsWriter.addElement("FieldData");
  {
    std::map<std::string, std::string> attribs = {{"type", "Float64"}, {"Name", "TimeValue"}, {"NumberOfTuples", "1"}};
    ptsWriter.addElement("DataArray", attribs);
    Vector<double> time_vec = {time};
    ptsWriter.writeDataArray(time_vec);
    ptsWriter.closeElement();
  }
  ptsWriter.closeElement();
  {
    std::map<std::string, std::string> attribs = {{"NumberOfPoints", std::to_string(n).c_str()},
                                                  {"NumberOfCells", std::to_string(n).c_str()}};
    ptsWriter.addElement("Piece", attribs);
  }
  auto arr = [&](const float* a) { return Vector<float>(a, a + n); };
  ptsWriter.addElement("Points");
  {
    std::map<std::string, std::string> attribs = {{"NumberOfComponents", "3"}, {"Name", "position"}, {"type", "Float32"}};
    ptsWriter.addElement("DataArray", attribs);
    std::array<Vector<float>, 3> xs = {arr(x), arr(y), arr(z)};
    Vector<float> pos = ptsWriter.unpackArray(xs);
    ptsWriter.writeDataArray(pos);
    ptsWriter.closeElement();
  }
  ptsWriter.closeElement();
  ptsWriter.addElement("Cells");
  {
    std::map<std::string, std::string> attribs = {{"Name", "connectivity"}, {"type", "Int32"}};
    ptsWriter.addElement("DataArray", attribs);
    Vector<int32_t> c(n);
    std::iota(c.begin(), c.end(), 0);
    ptsWriter.writeDataArray(c);
    ptsWriter.closeElement();
  }
  {
    std::map<std::string, std::string> attribs = {{"Name", "offsets"}, {"type", "Int32"}};
    ptsWriter.addElement("DataArray", attribs);
    Vector<int32_t> c(n);
    std::iota(c.begin(), c.end(), 1);
    ptsWriter.writeDataArray(c);
    ptsWriter.closeElement();
  }
  {
    std::map<std::string, std::string> attribs = {{"Name", "types"}, {"type", "UInt8"}};
    ptsWriter.addElement("DataArray", attribs);
    Vector<uint8_t> c(n);
    std::fill(c.begin(), c.end(), 1);
    ptsWriter.writeDataArray(c);
    ptsWriter.closeElement();
  }
  ptsWriter.closeElement();
  {
    std::map<std::string, std::string> attribs = {{"Vectors", "velocity"}};
    std::string scalar_list;
    if (has_radii) scalar_list.append("radius,");
    if (scalar_list.size() > 1) {
      scalar_list.pop_back();
      attribs.insert({"Scalars", scalar_list});
    }
    ptsWriter.addElement("PointData", attribs);
  }
  if (has_strengths) {
    std::map<std::string, std::string> attribs = {{"NumberOfComponents", "3"}, {"Name", "circulation"}, {"type", "Float32"}};
    ptsWriter.addElement("DataArray", attribs);
    std::array<Vector<float>, 3> ss = {arr(sx), arr(sy), arr(sz)};
    Vector<float> circ = ptsWriter.unpackArray(ss);
    ptsWriter.writeDataArray(circ);
    ptsWriter.closeElement();
  }
  if (has_radii) {
    std::map<std::string, std::string> attribs = {{"Name", "radius"}, {"type", "Float32"}};
    ptsWriter.addElement("DataArray", attribs);
    ptsWriter.writeDataArray(arr(r));
    ptsWriter.closeElement();
  }
  {
    std::map<std::string, std::string> attribs = {{"NumberOfComponents", "3"}, {"Name", "velocity"}, {"type", "Float32"}};
    ptsWriter.addElement("DataArray", attribs);
    std::array<Vector<float>, 3> us = {arr(u), arr(v), arr(w)};
    Vector<float> vel = ptsWriter.unpackArray(us);
    ptsWriter.writeDataArray(vel);
    ptsWriter.closeElement();
  }
  ptsWriter.closeElement();   // PointData
  ptsWriter.closeElement();   // Piece
  ptsWriter.finish();
  return 0;
}

// The call sequence of Surfaces::write_vtk (Surfaces.h:1301-1450): nodes, triangle
// cells (connectivity, offsets 3, 6, ..., type 5), and for an active or reactive
// surface the cell data "vortex sheet strength" (vs: the per-panel vector that
// vortex_sheet_to_vector makes) and, when the surface has them, "source sheet
// strength" (ss).  vs == nullptr: an inert surface, no cell data.
extern "C" int ref_write_panels_vtk(const char* fn, int64_t nn, const float* x, const float* y, const float* z,
                                    int64_t np, const uint32_t* idx, const float* vx, const float* vy,
                                    const float* vz, const float* ss, double time) {
  const bool has_vort_str = vx != nullptr, has_src_str = has_vort_str && ss != nullptr;
  VtkXmlWriter panelWriter = VtkXmlWriter(fn, true);
  panelWriter.addElement("FieldData");
  {
    std::map<std::string, std::string> attribs = {{"type", "Float64"}, {"Name", "TimeValue"}, {"NumberOfTuples", "1"}};
    panelWriter.addElement("DataArray", attribs);
    Vector<double> time_vec = {time};
    panelWriter.writeDataArray(time_vec);
    panelWriter.closeElement();
  }
  panelWriter.closeElement();
  {
    std::map<std::string, std::string> attribs = {{"NumberOfPoints", std::to_string(nn).c_str()},
                                                  {"NumberOfCells", std::to_string(np).c_str()}};
    panelWriter.addElement("Piece", attribs);
  }
  panelWriter.addElement("Points");
  {
    std::map<std::string, std::string> attribs = {{"NumberOfComponents", "3"}, {"Name", "position"}, {"type", "Float32"}};
    panelWriter.addElement("DataArray", attribs);
    std::array<Vector<float>, 3> xs = {Vector<float>(x, x + nn), Vector<float>(y, y + nn), Vector<float>(z, z + nn)};
    Vector<float> pos = panelWriter.unpackArray(xs);
    panelWriter.writeDataArray(pos);
    panelWriter.closeElement();
  }
  panelWriter.closeElement();
  panelWriter.addElement("Cells");
  {
    std::map<std::string, std::string> attribs = {{"Name", "connectivity"}, {"type", "Int32"}};
    panelWriter.addElement("DataArray", attribs);
    Vector<int32_t> v(idx, idx + 3 * np);
    panelWriter.writeDataArray(v);
    panelWriter.closeElement();
  }
  {
    std::map<std::string, std::string> attribs = {{"Name", "offsets"}, {"type", "Int32"}};
    panelWriter.addElement("DataArray", attribs);
    Vector<int32_t> v(np);
    std::iota(v.begin(), v.end(), 1);
    for (auto& o : v) o *= 3;
    panelWriter.writeDataArray(v);
    panelWriter.closeElement();
  }
  {
    std::map<std::string, std::string> attribs = {{"Name", "types"}, {"type", "UInt8"}};
    panelWriter.addElement("DataArray", attribs);
    Vector<uint8_t> v(np);
    std::fill(v.begin(), v.end(), 5);
    panelWriter.writeDataArray(v);
    panelWriter.closeElement();
  }
  panelWriter.closeElement();
  std::string scalar_list, vector_list;
  if (has_vort_str) vector_list.append("vortex sheet strength,");
  if (has_src_str) scalar_list.append("source sheet strength,");
  if (scalar_list.size() + vector_list.size() > 0) {
    {
      std::map<std::string, std::string> attribs;
      if (scalar_list.size() > 1) {
        scalar_list.pop_back();
        attribs.insert({"Scalars", scalar_list});
      }
      panelWriter.addElement("CellData", attribs);
    }
    if (has_vort_str) {
      std::map<std::string, std::string> attribs = {{"NumberOfComponents", "3"}, {"Name", "vortex sheet strength"},
                                                    {"type", "Float32"}};
      panelWriter.addElement("DataArray", attribs);
      std::array<Vector<float>, 3> vs = {Vector<float>(vx, vx + np), Vector<float>(vy, vy + np),
                                         Vector<float>(vz, vz + np)};
      Vector<float> strs = panelWriter.unpackArray(vs);
      panelWriter.writeDataArray(strs);
      panelWriter.closeElement();
    }
    if (has_src_str) {
      std::map<std::string, std::string> attribs = {{"Name", "source sheet strength"}, {"type", "Float32"}};
      panelWriter.addElement("DataArray", attribs);
      panelWriter.writeDataArray(Vector<float>(ss, ss + np));
      panelWriter.closeElement();
    }
    panelWriter.closeElement();   // CellData
  }
  panelWriter.closeElement();     // Piece
  panelWriter.finish();
  return 0;
}
