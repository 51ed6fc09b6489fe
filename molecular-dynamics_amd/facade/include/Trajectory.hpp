#pragma once

#include "System.hpp"

#include <cstdio>

// One extended-XYZ frame in particle-id order (reference: src/Trajectory.cpp:4-25)
void writeXyzFrame(std::FILE* file, const System& system, double time);
