// Structs.h (3D) -- point and segment types of the 3D drop-in API.
// Same members, order and defaults as the reference's 3D/Structs.h:15-36.
// Like the reference, the 2D and 3D SSegment types are separate APIs: include
// the headers of one of them per translation unit.
#pragma once

#include <vector>

struct point3D
{
    short x;
    short y;
    short z;
};

struct point3Df
{
    float x;
    float y;
    float z;
};

struct SSegment
{
    std::vector<point3D> points;
    bool pruned;
    bool is_end_segment;

    SSegment() : pruned(false), is_end_segment(false) {}
};
