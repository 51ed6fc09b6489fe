#include "Trajectory.hpp"

#include <algorithm>
#include <condition_variable>
#include <deque>
#include <exception>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "MdgpuRuntime.hpp"
#include "XyzFormat.hpp"

namespace {

void writeHeader(std::FILE* file, std::size_t n, double boxLength, double time) {
    std::fprintf(file, "%zu\n", n);
    std::fprintf(file,
                 "Lattice=\"%.10g 0 0 0 %.10g 0 0 0 %.10g\" Properties=species:S:1:pos:R:3:velo:R:3:id:I:1 "
                 "Time=%.6f pbc=\"T T T\"\n",
                 boxLength, boxLength, boxLength, time);
}

// The particle line of writeXyzFrame for id `id`, appended to out: formatted directly when every value is below
// 1e10 in magnitude, through snprintf (any length) otherwise.
void appendLine(std::string& out, const double* r, const double* v, std::size_t id) {
    bool direct = true;
    for (int k = 0; k < 3; ++k) direct = direct && xyz_format::fixed8_fits(r[k]) && xyz_format::fixed8_fits(v[k]);
    if (direct) {
        char line[xyz_format::particle_line_max];
        const char* end = xyz_format::particle_line(line, r, v, id);
        out.append(line, static_cast<std::size_t>(end - line));
        return;
    }
    static const char* format = "Ar %.8f %.8f %.8f %.8f %.8f %.8f %zu\n";
    char line[512];
    const int len = std::snprintf(line, sizeof line, format, r[0], r[1], r[2], v[0], v[1], v[2], id);
    if (len < static_cast<int>(sizeof line)) {
        out.append(line, static_cast<std::size_t>(len));
        return;
    }
    std::string longer(static_cast<std::size_t>(len) + 1, '\0');  // values up to 1.8e308 print 309 digits
    std::snprintf(longer.data(), longer.size(), format, r[0], r[1], r[2], v[0], v[1], v[2], id);
    out.append(longer.data(), static_cast<std::size_t>(len));
}

// Lines of a frame held as 6 doubles per id (x y z vx vy vz), formatted in parallel chunks for large frames and
// written in id order.
void writeFrameLines(std::FILE* file, const double* frame, std::size_t n) {
    constexpr std::size_t chunkMin = 65536;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const std::size_t chunks = std::min<std::size_t>({std::size_t(16), hw, (n + chunkMin - 1) / chunkMin});
    std::vector<std::string> text(std::max<std::size_t>(chunks, 1));
    auto format = [&](std::size_t c) {
        const std::size_t begin = n * c / chunks, end = n * (c + 1) / chunks;
        text[c].reserve((end - begin) * 96);
        for (std::size_t id = begin; id < end; ++id) appendLine(text[c], frame + 6 * id, frame + 6 * id + 3, id);
    };
    if (chunks <= 1) {
        if (n) {
            text[0].reserve(n * 96);
            for (std::size_t id = 0; id < n; ++id) appendLine(text[0], frame + 6 * id, frame + 6 * id + 3, id);
        }
    } else {
        std::vector<std::thread> pool;
        for (std::size_t c = 1; c < chunks; ++c) pool.emplace_back(format, c);
        format(0);
        for (auto& t : pool) t.join();
    }
    for (const auto& t : text) std::fwrite(t.data(), 1, t.size(), file);
}

}  // namespace

void writeXyzFrame(std::FILE* file, const System& system, double time) {
    const std::size_t n = system.size();
    // storage index of each id: the inverse of the cell re-sort permutation
    const auto ids = system.ids();
    std::vector<std::size_t> slot(n);
    for (std::size_t i = 0; i < n; ++i) slot[ids[i]] = i;
    std::vector<double> frame(6 * n);
    for (std::size_t id = 0; id < n; ++id) {
        const auto r = system.positionOf(slot[id]);
        const auto v = system.velocityOf(slot[id]);
        std::copy(r.begin(), r.end(), frame.begin() + 6 * id);
        std::copy(v.begin(), v.end(), frame.begin() + 6 * id + 3);
    }
    writeHeader(file, n, system.getBoxLength(), time);
    writeFrameLines(file, frame.data(), n);
}

struct XyzWriter::Impl {
    std::FILE* file;
    int slots;
    md_frames* frames = nullptr;
    std::size_t capacity = 0;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::pair<int, double>> queue;  // captured slots and their times, in call order
    int inFlight = 0;                          // captured and not yet printed
    bool stop = false;
    std::exception_ptr error;
    std::thread worker;

    Impl(std::FILE* f, int s) : file(f), slots(s) {
        worker = std::thread([this] { run(); });
    }

    void run() {
        for (;;) {
            std::pair<int, double> job;
            {
                std::unique_lock<std::mutex> lock(mu);
                cv.wait(lock, [&] { return stop || !queue.empty(); });
                if (queue.empty()) return;
                job = queue.front();
                queue.pop_front();
            }
            try {
                const double* frame = nullptr;
                std::size_t n = 0;
                double box = 0.0;
                mdgpu_facade::check(md_frames_wait(frames, job.first, &frame, &n, &box));
                writeHeader(file, n, box, job.second);
                writeFrameLines(file, frame, n);
            } catch (...) {
                std::lock_guard<std::mutex> lock(mu);
                if (!error) error = std::current_exception();
            }
            (void)md_frames_release(frames, job.first);
            {
                std::lock_guard<std::mutex> lock(mu);
                --inFlight;
            }
            cv.notify_all();
        }
    }

    void rethrow() {
        if (error) {
            auto e = error;
            error = nullptr;
            std::rethrow_exception(e);
        }
    }

    // Waits (lock held) until at most `limit` frames are in flight.
    void drain(std::unique_lock<std::mutex>& lock, int limit) {
        cv.wait(lock, [&] { return inFlight <= limit; });
    }
};

XyzWriter::XyzWriter(std::FILE* file, int slots) : impl_(std::make_unique<Impl>(file, std::max(1, slots))) {}

XyzWriter::~XyzWriter() {
    {
        std::unique_lock<std::mutex> lock(impl_->mu);
        impl_->drain(lock, 0);
        impl_->stop = true;
    }
    impl_->cv.notify_all();
    impl_->worker.join();
    std::fflush(impl_->file);
    if (impl_->frames) (void)md_frames_destroy(impl_->frames);
}

void XyzWriter::write(const System& system, double time) {
    md_state* state = system.deviceState();
    std::unique_lock<std::mutex> lock(impl_->mu);
    impl_->rethrow();
    if (!impl_->frames || system.size() > impl_->capacity) {
        impl_->drain(lock, 0);
        if (impl_->frames) mdgpu_facade::check(md_frames_destroy(impl_->frames));
        impl_->frames = nullptr;
        impl_->capacity = system.size();
        mdgpu_facade::check(md_frames_create(mdgpu_facade::context(), impl_->capacity, impl_->slots, &impl_->frames));
    }
    impl_->drain(lock, impl_->slots - 1);
    int slot = -1;
    mdgpu_facade::check(md_frames_capture(impl_->frames, state, &slot));
    ++impl_->inFlight;
    impl_->queue.emplace_back(slot, time);
    lock.unlock();
    impl_->cv.notify_all();
}

void XyzWriter::flush() {
    std::unique_lock<std::mutex> lock(impl_->mu);
    impl_->drain(lock, 0);
    std::fflush(impl_->file);
    impl_->rethrow();
}
