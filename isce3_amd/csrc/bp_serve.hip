// Continuous-batching tile server over one resident scene
// (include/isce3_amd_backproject.h, "serving").
//
// Submitters append requests to a FIFO and get a ticket; one worker thread
// repeatedly takes the longest prefix of the FIFO that fits the scene
// (max_pixels, max_blocks), focuses it with one bp_scene_focus call -- one
// geometry and one coherent-sum launch for the whole batch, replayed from the
// scene's HIP graph -- on the server's own stream, and completes the
// tickets.  Requests that arrive while a batch runs form the next batch, so
// the batch size follows the load: one request when idle, many under load.
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/isce3_amd_backproject.h"
#include "bp_internal.h"

void bp_set_last_error(const std::string& msg);
int bp_host_exception() noexcept;

namespace {

int fail(int code, const std::string& msg)
{
    bp_set_last_error(msg);
    return code;
}

int64_t block_pixels(const bp_out_block& b) { return b.grid.length * b.grid.width; }

struct Result {
    bool done = false;
    int code = BP_SUCCESS;
    int failed = 0;
    std::string msg;   // the batch's error message when code < 0
};

}  // namespace

struct bp_server {
    bp_scene* scene = nullptr;
    bp_options options{};
    int max_pending = 0;
    int device = 0;
    int64_t max_pixels = 0;
    int max_blocks = 0;
    hipStream_t stream = nullptr;

    std::mutex mu;
    std::condition_variable work_cv, done_cv;
    std::deque<std::pair<uint64_t, bp_out_block>> pending;
    std::unordered_map<uint64_t, Result> results;
    uint64_t next_ticket = 1;
    bool stopping = false;
    int waiters = 0;                  // threads inside bp_serve_wait
    std::condition_variable left_cv;  // a waiter left (destroy waits for none)
    std::thread worker;

    void run()
    {
        (void)hipSetDevice(device);
        std::vector<bp_out_block> batch;
        std::vector<uint64_t> tickets;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu);
                work_cv.wait(lk, [&] { return stopping || !pending.empty(); });
                if (pending.empty()) return;  // stopping and drained
                batch.clear();
                tickets.clear();
                int64_t npix = 0;
                while (!pending.empty() && (int)batch.size() < max_blocks) {
                    const int64_t p = block_pixels(pending.front().second);
                    if (!batch.empty() && npix + p > max_pixels) break;
                    npix += p;
                    tickets.push_back(pending.front().first);
                    batch.push_back(pending.front().second);
                    pending.pop_front();
                }
            }
            int any = 0;
            const int code = bp_scene_focus(scene, batch.data(), (int)batch.size(), &options,
                                            stream, &any, nullptr);
            const std::string msg = code < 0 ? std::string(bp_last_error_message()) : std::string();
            {
                std::lock_guard<std::mutex> lk(mu);
                for (size_t i = 0; i < batch.size(); ++i) {
                    Result& r = results[tickets[i]];
                    r.done = true;
                    r.code = code < 0 ? code : (batch[i].failed ? BP_FAILED_TO_CONVERGE : BP_SUCCESS);
                    r.failed = batch[i].failed;
                    r.msg = msg;
                }
            }
            done_cv.notify_all();
        }
    }
};

extern "C" {

int bp_server_create(bp_scene* scene, const bp_options* options, int max_pending,
                     bp_server** server)
try {
    if (server == nullptr || scene == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    *server = nullptr;
    if (max_pending < 1) return fail(BP_ERR_INVALID_ARGUMENT, "max_pending must be > 0");
    auto* s = new bp_server();
    s->scene = scene;
    if (options) s->options = *options;
    s->max_pending = max_pending;
    s->device = bp::scene_device(scene);
    s->max_pixels = bp::scene_max_pixels(scene);
    s->max_blocks = bp::scene_max_blocks(scene);
    if (hipSetDevice(s->device) != hipSuccess ||
        hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return fail(BP_ERR_HIP, "server stream creation failed");
    }
    try {
        s->worker = std::thread([s] { s->run(); });
    } catch (...) {
        (void)hipStreamDestroy(s->stream);
        delete s;
        return fail(BP_ERR_RUNTIME, "server thread creation failed");
    }
    *server = s;
    return BP_SUCCESS;
} catch (...) {
    return bp_host_exception();
}

int bp_server_destroy(bp_server* s)
try {
    if (s == nullptr) return BP_SUCCESS;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->work_cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    // every accepted request is complete now; let the threads still inside
    // bp_serve_wait return before the server goes away
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->done_cv.notify_all();
        s->left_cv.wait(lk, [&] { return s->waiters == 0; });
    }
    (void)hipSetDevice(s->device);
    (void)hipStreamDestroy(s->stream);
    delete s;
    return BP_SUCCESS;
} catch (...) {
    return bp_host_exception();
}

int bp_serve_submit(bp_server* s, const bp_out_block* block, uint64_t* ticket)
try {
    if (s == nullptr || block == nullptr || ticket == nullptr)
        return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    // the per-block rules of the batched call, checked here so that a bad
    // request is refused alone instead of failing the batch it would join
    const int bad = bp::check_out_block(*block);
    if (bad != BP_SUCCESS) return bad;
    if (block_pixels(*block) > s->max_pixels)
        return fail(BP_ERR_LENGTH, "request of " + std::to_string(block_pixels(*block)) +
                                       " pixels exceeds the scene's max_pixels");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping) return fail(BP_ERR_RUNTIME, "server is shutting down");
        if ((int)s->pending.size() >= s->max_pending)
            return fail(BP_ERR_LENGTH, "serving queue is full");
        *ticket = s->next_ticket++;
        bp_out_block b = *block;
        b.failed = 0;
        s->pending.emplace_back(*ticket, b);
        s->results[*ticket] = Result{};
    }
    s->work_cv.notify_one();
    return BP_SUCCESS;
} catch (...) {
    return bp_host_exception();
}

int bp_serve_poll(bp_server* s, uint64_t ticket, int* done, int* failed)
try {
    if (s == nullptr || done == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->results.find(ticket);
    if (it == s->results.end()) return fail(BP_ERR_INVALID_ARGUMENT, "unknown ticket");
    *done = it->second.done ? 1 : 0;
    if (!it->second.done) return BP_SUCCESS;
    if (failed) *failed = it->second.failed;
    const int code = it->second.code;
    const std::string msg = it->second.msg;
    s->results.erase(it);
    if (code < 0) return fail(code, "focusing the request's batch failed: " + msg);
    return code;
} catch (...) {
    return bp_host_exception();
}

int bp_serve_wait(bp_server* s, uint64_t ticket, int* failed)
try {
    if (s == nullptr) return fail(BP_ERR_INVALID_ARGUMENT, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->results.find(ticket);
    if (it == s->results.end()) return fail(BP_ERR_INVALID_ARGUMENT, "unknown ticket");
    // A ticket has one consumer: if another thread's wait or poll reports it
    // first, the entry disappears and this wait returns "unknown ticket".
    ++s->waiters;
    s->done_cv.wait(lk, [&] {
        auto f = s->results.find(ticket);
        return f == s->results.end() || f->second.done;
    });
    // destroy, woken here, still needs the mutex this thread holds until it
    // no longer touches the server
    if (--s->waiters == 0) s->left_cv.notify_all();
    it = s->results.find(ticket);
    int code = BP_ERR_INVALID_ARGUMENT;
    std::string msg = "unknown ticket (reported to another caller)";
    if (it != s->results.end()) {
        if (failed) *failed = it->second.failed;
        code = it->second.code;
        msg = "focusing the request's batch failed: " + it->second.msg;
        s->results.erase(it);
    }
    lk.unlock();
    if (code < 0) return fail(code, msg);
    return code;
} catch (...) {
    return bp_host_exception();
}

}  // extern "C"
