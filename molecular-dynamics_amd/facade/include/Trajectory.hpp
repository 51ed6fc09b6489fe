#pragma once

#include "System.hpp"

#include <cstdio>
#include <memory>

// One extended-XYZ frame in particle-id order (reference: src/Trajectory.cpp:4-25)
void writeXyzFrame(std::FILE* file, const System& system, double time);

// ---- mdgpu extension ----------------------------------------------------------------------------------------
// The same frames written asynchronously. write() captures the System's positions and velocities on the GPU, in
// id order, as they stand at the call, and returns at once; the download overlaps the steps that follow, and a
// writer thread prints each frame byte for byte as writeXyzFrame does, in call order. With every slot in flight,
// write() waits for the oldest frame to be printed. flush() returns once every frame written so far is in the file
// (fflush'ed); the destructor flushes. The FILE stays the caller's and must stay open until then. A failure on the
// writer thread is rethrown by the next write() or flush().
class XyzWriter {
public:
    explicit XyzWriter(std::FILE* file, int slots = 3);
    ~XyzWriter();
    XyzWriter(const XyzWriter&) = delete;
    XyzWriter& operator=(const XyzWriter&) = delete;

    void write(const System& system, double time);
    void flush();

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};
