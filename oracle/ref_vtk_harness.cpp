// ref_vtk_harness.cpp -- test infrastructure only.  Drives the reference's own VTK
// XML writer (VtkXmlWriter.h with its vendored tinyxml2 and cppcodec, read in place
// from the reference checkout) through the call sequence of Points::write_vtk
// (Points.h:851-1039).  The sequence is restated here because Points.h needs Eigen3
// (through ElementBase.h -> Body.h).  Built by build_ref.sh into
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
