// image.hpp -- Image, the drop-in for src/Simulation/Image/image.hpp:21-54:
// the intensities leaving every cell's ray pair along one ray direction and
// the cells' coordinates on the image plane.
#pragma once

#include <string>

#include "io.hpp"
#include "model.hpp"

namespace magritte {

struct Image {
    const long ray_nr;   // number of the ray to be imaged
    Double1 ImX, ImY;    // coordinates of each cell on the image plane
    Double2 I_p, I_m;    // [ncells][nfreqs] intensities out along / against the ray

    Image(long ray_nr, const Parameters &parameters);

    // Image::write (image.cpp:48-86): Simulation/Image/{ImX,ImY,I_m,I_p}_<ray_nr>
    void write(const Io &io) const;
    // Image::set_coordinates (image.cpp:93-132)
    void set_coordinates(const Geometry &geometry);

  private:
    long ncells, nfreqs;
    static const std::string prefix;
};

}  // namespace magritte
