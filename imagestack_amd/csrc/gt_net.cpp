// TCP front end that speaks ImageStack's own -send / -receive protocol
// (TCPConnection::sendImage / recvImage, src/Network.cpp:161-196, driven by
// Send::apply and Receive::apply, src/NetworkOps.cpp:79-121): a header of
// four uint32 (width, height, frames, channels) and then the float32 samples
// scanline by scanline in channel, frame, row order (the planar layout used
// throughout this library).  One image per connection.
//
// Each connection gets a handler thread that reads the image, runs the
// configured operation, and sends the result as the same kind of message:
// back on the same connection (reply_port 0), or - so that an unmodified
// reference client can run  ImageStack -load a.tmp -send HOST P -receive R
// -save b.tmp  - on a new connection to the sender's address at reply_port,
// retried until the client's -receive is listening.  Bilateral requests go
// through the serving loop (gt_server: HBM admission, continuous batching);
// the shapes it does not take (several frames, channel counts other than 1
// and 3) run on the front end's own context.
//
// Headers are unauthenticated: an image is accepted only within max_pixels
// and max_bytes, and only while the bytes every handler holds (input and
// output) stay within max_inflight; a request past that is refused, its
// connection closed, before anything is allocated.  The listener binds to
// loopback unless told otherwise.
#include <arpa/inet.h>
#include <netinet/in.h>
#include <netinet/tcp.h>
#include <poll.h>
#include <sys/socket.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <set>
#include <thread>
#include <vector>

#include "gt_internal.h"
#include "../../include/gtlattice.h"

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

bool recv_all(int fd, void *buf, size_t len) {
    char *p = (char *)buf;
    while (len > 0) {
        const ssize_t r = ::recv(fd, p, len, 0);
        if (r <= 0) return false;
        p += r;
        len -= (size_t)r;
    }
    return true;
}

bool send_all(int fd, const void *buf, size_t len) {
    const char *p = (const char *)buf;
    while (len > 0) {
        const ssize_t r = ::send(fd, p, len, MSG_NOSIGNAL);
        if (r <= 0) return false;
        p += r;
        len -= (size_t)r;
    }
    return true;
}

bool send_image(int fd, const uint32_t hdr[4], const std::vector<float> &px) {
    return send_all(fd, hdr, 16) && send_all(fd, px.data(), px.size() * sizeof(float));
}

} // namespace

struct gt_net_server {
    gt_net_config cfg{};
    int device = 0;
    int listen_fd = -1;
    int port = 0;
    std::atomic<bool> stop{false};
    std::thread acceptor;
    std::mutex mu;                    // guards the fields below
    std::condition_variable cv;
    std::set<int> open_fds;
    int active = 0;                   // detached handler threads still running
    long long received = 0, replied = 0, failed = 0, refused = 0;
    long long inflight = 0;           // bytes reserved by handlers (input + output images)
    std::vector<double> lat_ms;
    gt_server *serving = nullptr;     // bilateral requests of 1 or 3 channels, 1 frame
    gt_ctx *ctx = nullptr;            // the other shapes
    std::mutex ctx_mu;
};

namespace {

constexpr int kMaxHandlers = 64;
constexpr size_t kLatWindow = 100000;

int connect_back(gt_net_server *s, const sockaddr_in &peer, int port, double give_up_ms) {
    sockaddr_in a = peer;
    a.sin_port = htons((uint16_t)port);
    const double end = now_ms() + give_up_ms;
    while (!s->stop.load()) {
        const int fd = ::socket(AF_INET, SOCK_STREAM, IPPROTO_TCP);
        if (fd < 0) return -1;
        if (::connect(fd, (const sockaddr *)&a, sizeof(a)) == 0) return fd;
        ::close(fd);
        if (now_ms() > end) return -1;
        std::this_thread::sleep_for(std::chrono::milliseconds(10));
    }
    return -1;
}

// Filters one received image; returns GT_OK or an error code.
int run_op(gt_net_server *s, const uint32_t h[4], std::vector<float> &im, std::vector<float> &out) {
    const int W = (int)h[0], H = (int)h[1], F = (int)h[2], C = (int)h[3];
    if (s->cfg.op == GT_NET_ECHO) {
        out.swap(im);
        return GT_OK;
    }
    out.assign(im.size(), 0.0f);
    if (F == 1 && (C == 1 || C == 3)) {
        gt_request r{};
        r.W = W;
        r.H = H;
        r.C = C;
        r.im = im.data();
        r.row_pitch = W;
        r.plane_pitch = (long long)W * H;
        r.out = out.data();
        r.sigma_s = s->cfg.sigma_s;
        r.sigma_r = s->cfg.sigma_r;
        r.storage = s->cfg.storage;
        long long ticket = 0;
        int rc = gt_server_submit(s->serving, &r, &ticket);
        if (rc != GT_OK) return rc;
        gt_request_result res{};
        return gt_server_wait(s->serving, ticket, &res);
    }
    std::lock_guard<std::mutex> lk(s->ctx_mu);
    return gt_joint_bilateral_host(s->ctx, W, H, F, C, im.data(), C, im.data(), s->cfg.sigma_s, s->cfg.sigma_s, 0.0f,
                                   s->cfg.sigma_r, out.data(), s->cfg.storage, nullptr, nullptr);
}

void handle(gt_net_server *s, int fd, sockaddr_in peer) {
    bool ok = false, refused = false;
    long long held = 0;   // bytes reserved against max_inflight
    double t0 = 0, t1 = 0;
    uint32_t h[4];
    if (recv_all(fd, h, sizeof(h))) {
        const uint64_t n = (uint64_t)h[0] * h[1] * h[2];
        const uint64_t bytes = n * h[3] * sizeof(float);
        bool sane = h[0] >= 1 && h[1] >= 1 && h[2] >= 1 && h[3] >= 1 && h[3] < GT_MAX_VD &&
                    n <= (uint64_t)s->cfg.max_pixels && bytes <= (uint64_t)s->cfg.max_bytes;
        if (sane) {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->inflight + 2 * (long long)bytes > s->cfg.max_inflight) {
                refused = true;
                sane = false;
            } else {
                held = 2 * (long long)bytes;
                s->inflight += held;
            }
        }
        std::vector<float> im, out;
        if (sane) {
            im.resize(n * h[3]);
            if (recv_all(fd, im.data(), im.size() * sizeof(float))) {
                {
                    std::lock_guard<std::mutex> lk(s->mu);
                    s->received++;
                }
                t0 = now_ms();
                if (run_op(s, h, im, out) == GT_OK) {
                    if (s->cfg.reply_port > 0) {
                        {
                            std::lock_guard<std::mutex> lk(s->mu);
                            s->open_fds.erase(fd);
                        }
                        ::close(fd);
                        fd = connect_back(s, peer, s->cfg.reply_port, 10000.0);
                        if (fd >= 0) {
                            std::lock_guard<std::mutex> lk(s->mu);
                            s->open_fds.insert(fd);
                        }
                    }
                    ok = fd >= 0 && send_image(fd, h, out);
                    t1 = now_ms();
                }
            }
        }
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (fd >= 0) {
        s->open_fds.erase(fd);
        ::close(fd);
    }
    s->inflight -= held;
    if (refused) {
        s->refused++;
    } else if (ok) {
        s->replied++;
        // latencies of the last kLatWindow replies (a ring), for the percentiles
        if (s->lat_ms.size() < kLatWindow) s->lat_ms.push_back(t1 - t0);
        else s->lat_ms[(size_t)(s->replied - 1) % kLatWindow] = t1 - t0;
    } else {
        s->failed++;
    }
    s->active--;
    s->cv.notify_all();   // under the lock: the destroyer frees s only after this thread unlocks
}

void accept_loop(gt_net_server *s) {
    while (!s->stop.load()) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->active < kMaxHandlers || s->stop.load(); });
        }
        pollfd p{s->listen_fd, POLLIN, 0};
        if (::poll(&p, 1, 100) <= 0) continue;
        sockaddr_in peer{};
        socklen_t len = sizeof(peer);
        const int fd = ::accept(s->listen_fd, (sockaddr *)&peer, &len);
        if (fd < 0) continue;
        timeval tv{30, 0};   // a client that stops sending is dropped after 30 s
        ::setsockopt(fd, SOL_SOCKET, SO_RCVTIMEO, &tv, sizeof(tv));
        const int one = 1;
        ::setsockopt(fd, IPPROTO_TCP, TCP_NODELAY, &one, sizeof(one));
        std::lock_guard<std::mutex> lk(s->mu);
        s->active++;
        s->open_fds.insert(fd);
        std::thread(handle, s, fd, peer).detach();
    }
}

} // namespace

extern "C" {

int gt_net_server_create(gt_net_server **out, int device, const gt_net_config *cfg) {
    if (!out || !cfg) return gt::set_error(GT_EINVAL, "null pointer");
    *out = nullptr;
    if (cfg->op != GT_NET_ECHO && cfg->op != GT_NET_BILATERAL) return gt::set_error(GT_EINVAL, "unknown operation");
    if (cfg->op == GT_NET_BILATERAL && !(cfg->sigma_s > 0 && cfg->sigma_r > 0))
        return gt::set_error(GT_EINVAL, "bilateral sigmas must be positive");
    if (cfg->port < 0 || cfg->port > 65535 || cfg->reply_port < 0 || cfg->reply_port > 65535)
        return gt::set_error(GT_EINVAL, "bad port");
    gt_net_server *s = new gt_net_server();
    s->cfg = *cfg;
    s->device = device;
    if (s->cfg.max_pixels <= 0) s->cfg.max_pixels = 1ll << 26;
    if (s->cfg.max_bytes <= 0) s->cfg.max_bytes = 1ll << 30;
    if (s->cfg.max_inflight <= 0) s->cfg.max_inflight = 4ll << 30;
    in_addr bind_to{};
    const char *addr = (cfg->bind_addr && cfg->bind_addr[0]) ? cfg->bind_addr : "127.0.0.1";
    s->cfg.bind_addr = nullptr;   // the caller's string is not kept
    if (::inet_pton(AF_INET, addr, &bind_to) != 1) {
        delete s;
        return gt::set_error(GT_EINVAL, "bind address is not an IPv4 address");
    }
    if (cfg->op == GT_NET_BILATERAL) {
        int rc = gt_server_create(&s->serving, device, &s->cfg.serving);
        if (rc == GT_OK) rc = gt_ctx_create(&s->ctx, device);
        if (rc != GT_OK) {
            gt_net_server_destroy(s);
            return rc;
        }
    }
    s->listen_fd = ::socket(AF_INET, SOCK_STREAM, IPPROTO_TCP);
    const int one = 1;
    sockaddr_in a{};
    a.sin_family = AF_INET;
    a.sin_addr = bind_to;
    a.sin_port = htons((uint16_t)cfg->port);
    socklen_t len = sizeof(a);
    if (s->listen_fd < 0 || ::setsockopt(s->listen_fd, SOL_SOCKET, SO_REUSEADDR, &one, sizeof(one)) != 0 ||
        ::bind(s->listen_fd, (sockaddr *)&a, sizeof(a)) != 0 || ::listen(s->listen_fd, 64) != 0 ||
        ::getsockname(s->listen_fd, (sockaddr *)&a, &len) != 0) {
        gt_net_server_destroy(s);
        return gt::set_error(GT_EINVAL, "could not listen on the port");
    }
    s->port = ntohs(a.sin_port);
    s->acceptor = std::thread(accept_loop, s);
    *out = s;
    return GT_OK;
}

int gt_net_server_port(gt_net_server *s) { return s ? s->port : -1; }

int gt_net_server_stats(gt_net_server *s, gt_net_stats *st) {
    if (!s || !st) return gt::set_error(GT_EINVAL, "null pointer");
    std::lock_guard<std::mutex> lk(s->mu);
    memset(st, 0, sizeof(*st));
    st->received = s->received;
    st->replied = s->replied;
    st->failed = s->failed;
    st->refused = s->refused;
    st->inflight_bytes = s->inflight;
    std::vector<double> v = s->lat_ms;
    std::sort(v.begin(), v.end());
    if (!v.empty()) {
        st->p50_ms = v[(v.size() - 1) / 2];
        st->p99_ms = v[std::min(v.size() - 1, (size_t)(0.99 * (double)v.size()))];
    }
    return GT_OK;
}

void gt_net_server_destroy(gt_net_server *s) {
    if (!s) return;
    s->stop.store(true);
    s->cv.notify_all();
    if (s->acceptor.joinable()) s->acceptor.join();
    {
        std::unique_lock<std::mutex> lk(s->mu);
        for (int fd : s->open_fds) ::shutdown(fd, SHUT_RDWR);
        s->cv.wait(lk, [&] { return s->active == 0; });
    }
    if (s->listen_fd >= 0) ::close(s->listen_fd);
    if (s->serving) gt_server_destroy(s->serving);
    if (s->ctx) gt_ctx_destroy(s->ctx);
    delete s;
}

} // extern "C"
