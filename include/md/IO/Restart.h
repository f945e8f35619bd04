// Amber restart (rst7 / inpcrd) files: a reader that does not depend on blank
// separators, knows velocities, and a writer.
//
// The reference reads an rst7 inside AmberTopology (src/AmberTopology/
// AmberTopology.cpp:388-438): it skips the title and atom-count lines, splits
// the rest on blanks, takes the first 3N values as coordinates and any six
// after them as the box, so velocities are read as the box and touching
// fields break the split.  It has no writer.  Semantics::REF keeps that
// behaviour; this reader serves Semantics::AMBER and any caller that needs
// velocities or writes restarts.
#ifndef MD_AMD_IO_RESTART_H
#define MD_AMD_IO_RESTART_H

#include <array>
#include <string>
#include <vector>

namespace md::io {

struct Restart {
    std::string title;
    int n_atoms = 0;
    bool has_time = false;
    double time = 0.0;               // ps, from the atom-count line when present
    std::vector<double> coords;      // 3N, A
    std::vector<double> velocities;  // 3N, A/ps (Amber's file unit is A/(20.455 ps); converted) or empty
    bool has_box = false;
    std::array<double, 6> box{0, 0, 0, 90, 90, 90};  // lengths (A) and angles (degrees)
};

// Amber velocity unit in a restart file: A per 1/20.455 ps.
constexpr double kRestartVelocityScale = 20.455;

// Reads title, atom count (and time), 3N coordinates, optional 3N velocities
// and an optional box (3 lengths, or 3 lengths and 3 angles).  Numbers are
// taken from 12-column fields when every field of a line parses (so touching
// fields such as "-123.4567890-12.3456789" are read), otherwise from the
// blank-separated tokens.  Throws std::runtime_error on an unreadable file or
// a value count that fits none of the layouts.
Restart read_rst7(const std::string& path);

// Writes the Amber ASCII restart layout: title (A80), atom count and time
// (I6, E15.7), coordinates and velocities in 6F12.7, the box in 6F12.7.
// Velocities are converted to the file unit.  Throws std::runtime_error when
// the file cannot be written or a value does not fit F12.7.
void write_rst7(const std::string& path, const Restart& r);

}  // namespace md::io

#endif
