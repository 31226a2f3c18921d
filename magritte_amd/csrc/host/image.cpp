// image.cpp -- Image members (src/Simulation/Image/image.cpp).
#include "image.hpp"

#include <cmath>

namespace magritte {

const std::string Image::prefix = "Simulation/Image/";

Image::Image(long ray_nr_, const Parameters &parameters)
    : ray_nr(ray_nr_), ncells((long)parameters.ncells()), nfreqs((long)parameters.nfreqs())
{
    ImX.assign(ncells, 0.0);
    ImY.assign(ncells, 0.0);
    I_p.assign(ncells, Double1(nfreqs, 0.0));
    I_m.assign(ncells, Double1(nfreqs, 0.0));
}

void Image::write(const Io &io) const
{
    const std::string nr = std::to_string(ray_nr);
    io.write_list(prefix + "ImX_" + nr, ImX);
    io.write_list(prefix + "ImY_" + nr, ImY);
    io.write_array(prefix + "I_m_" + nr, I_m);
    io.write_array(prefix + "I_p_" + nr, I_p);
}

// Unit vectors of the image plane: i = (ry, -rx, 0)/d and
// j = (rx rz, ry rz, -d)/d with d = sqrt(rx^2 + ry^2); a ray along z keeps
// the x and y coordinates.
void Image::set_coordinates(const Geometry &geometry)
{
    const Vector3d &r = geometry.rays.rays.at(ray_nr);
    const double rx = r[0], ry = r[1], rz = r[2];
    const double denominator = std::sqrt(rx * rx + ry * ry);
    const double inverse_denominator = 1.0 / denominator;
    const double ix = ry * inverse_denominator;
    const double iy = -rx * inverse_denominator;
    const double jx = rx * rz * inverse_denominator;
    const double jy = ry * rz * inverse_denominator;
    const double jz = -denominator;
    const Vector3d1 &pos = geometry.cells.position;
    for (long p = 0; p < ncells; p++) {
        if (denominator != 0.0) {
            ImX[p] = ix * pos[p][0] + iy * pos[p][1];
            ImY[p] = jx * pos[p][0] + jy * pos[p][1] + jz * pos[p][2];
        } else {
            ImX[p] = pos[p][0];
            ImY[p] = pos[p][1];
        }
    }
}

}  // namespace magritte
