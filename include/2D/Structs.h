// Structs.h (2D) -- point and segment types of the 2D drop-in API.
// Same members, order and defaults as the reference's 2D/Structs.h:15-34, so
// code written against it compiles and behaves unchanged.
#pragma once

#include <vector>

struct point2D
{
    short x;
    short y;
};

struct point2Df
{
    float x;
    float y;
};

struct SSegment
{
    std::vector<point2D> points;
    bool pruned;
    bool is_end_segment;

    SSegment() : pruned(false), is_end_segment(false) {}
};
