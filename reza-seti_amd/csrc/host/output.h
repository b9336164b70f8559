// Output files of Simulation::WriteOutput (Simulation.cpp:835-856) written
// from host arrays: the fluid-variables VTU of OutputFile::WriteVTK_Collective
// (OutputFile.cpp:60-231), the near-wall CSV of
// OutputFile::WriteSurfaceAttribute (:234-316) and its smoothed copy of
// OutputFile::processFlowData / smoothData (:369-483).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "geometry.h"

namespace lbm {

// Points are the nodes' grid indices in z-y-x loop order; point data
// "Velocity" (3), "Density", "VelocityMagnitude"; one hexahedron (VTK order)
// per grid cube whose 8 corners are nodes, over the reference's cube range
// x < NX - 2, y < NY - 2, z < NZ - 2 (Geometry::nx counts intervals, and the
// loop stops one short of it). XML, appended raw binary, UInt64 headers.
bool write_vtu(const Tables& t, const std::vector<double>& rho, const std::vector<double>& u,
               const std::string& path, std::string* err);

// X,Y,Z,WSS[,OSI,TAWSS],VelMag per near-wall node (coordinates = grid point
// positions), written with the default stream format (6 significant digits).
bool write_surface_csv(const Tables& t, const std::vector<double>& wss, const std::vector<double>& osi,
                       const std::vector<double>& tawss, const std::vector<double>& velmag, bool pulsatile,
                       const std::string& path, std::string* err);

// Reads a surface CSV back (the values as printed) and writes it with WSS
// (and OSI, TAWSS when pulsatile) replaced by the mean over the other rows
// within `radius` (0 < distance <= radius), or kept where there are none.
bool smooth_surface_csv(const std::string& in_path, const std::string& out_path, double radius, bool pulsatile,
                        std::string* err);

}  // namespace lbm
