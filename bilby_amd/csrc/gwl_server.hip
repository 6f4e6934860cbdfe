// Continuous-batching request server (host code, one scheduler thread).
//
// Samplers submit requests (an event id, N parameter rows in host memory, a
// host output buffer); the scheduler keeps at most one batch in flight per
// unit (an event, a ROQ likelihood, or a group of events evaluated by one
// multi-event launch).  Whenever a unit is idle and has pending requests,
// all of them (up to max_rows rows per event) are packed into the unit's
// pinned staging buffer, copied to the device, evaluated on the unit's
// stream and copied back; completions are found by polling the units' HIP
// events, and each finished request's results are copied to its output
// buffer and its waiters are woken.  Callers block in gwl_server_wait
// without holding any interpreter lock; latency is measured here, from
// submission to the copy into the output buffer.
//
// bilby runs a pool of single-dict callers instead
// (bilby/core/sampler/base_sampler.py:772-800); this is the native side of
// bilby_amd.serving.
#include "gwl_internal.h"

#include <sys/socket.h>
#include <sys/stat.h>
#include <sys/un.h>
#include <unistd.h>

#include <cerrno>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

enum { UNIT_EVENT = 0, UNIT_ROQ = 1, UNIT_GROUP = 2, UNIT_RELBIN = 3, UNIT_CALIB = 4 };

struct Req {
  int id, unit, member;
  const double* rows;
  int64_t n;
  double* out;
  int64_t ticket;
  int64_t off;  // row offset in the member's staging buffer
  Clock::time_point t0;
};

struct Ticket {
  bool done = false;
  int status = GWL_OK;
  double latency = 0.0;
};

struct Unit {
  int kind = UNIT_EVENT;
  gwl_event* ev = nullptr;
  int precision = GWL_FP32;
  gwl_roq* roq = nullptr;
  int mode = 1;
  gwl_multi* multi = nullptr;
  gwl_relbin* rb = nullptr;
  gwl_calib* cal = nullptr;
  int n_members = 1;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  std::vector<double*> h_rows, h_out;  // pinned, per member
  std::vector<double*> d_rows, d_out;  // device, per member
  bool busy = false;
  std::vector<Req> inflight;
};

}  // namespace

struct gwl_server {
  gwl_ctx* ctx = nullptr;
  int64_t max_rows = 0;
  std::vector<Unit*> units;
  std::vector<std::pair<int, int>> ids;  // id -> (unit, member)
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req> pending;
  std::unordered_map<int64_t, Ticket> tickets;
  int64_t next_ticket = 1;
  bool stop = false;
  std::thread th;
  int64_t batches = 0, rows_done = 0;
  int64_t device_bytes = 0;  // staging and group buffers held by the server
  std::string last_error;
  // socket front-end (gwl_server_listen)
  int listen_fd = -1;
  std::string sock_path;
  std::thread accept_th;
  std::mutex conn_mu;
  std::vector<struct SockConn*> conns;
  bool closing = false;
};

struct SockConn {
  int fd = -1;
  std::thread th;
};

namespace {

void finish(gwl_server* s, Unit* u, int status) {
  const Clock::time_point t = Clock::now();
  std::lock_guard<std::mutex> lk(s->mu);
  int64_t rows = 0;
  for (const Req& r : u->inflight) {
    Ticket& tk = s->tickets[r.ticket];
    tk.done = true;
    tk.status = status;
    tk.latency = std::chrono::duration<double>(t - r.t0).count();
    rows += r.n;
  }
  s->batches += 1;
  s->rows_done += rows;
  u->inflight.clear();
  u->busy = false;
  s->cv_done.notify_all();
}

void fail(gwl_server* s, Unit* u, int status, const char* what) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    const char* msg = gwl_last_error();
    s->last_error = std::string(what) + ": " + (msg && *msg ? msg : "failed");
  }
  finish(s, u, status ? status : GWL_E_HIP);
}

// Pack, upload, evaluate and download one batch of a unit (asynchronously
// on its stream); the requests move to u->inflight.
void launch(gwl_server* s, Unit* u, std::vector<Req>& reqs) {
  std::vector<int64_t> rows(u->n_members, 0);
  for (Req& r : reqs) {
    r.off = rows[r.member];
    std::memcpy(u->h_rows[r.member] + r.off * GWL_NPARAM, r.rows, (size_t)r.n * GWL_NPARAM * 8);
    rows[r.member] += r.n;
  }
  u->inflight.swap(reqs);
  reqs.clear();
  u->busy = true;
  hipError_t err = hipSuccess;
  for (int m = 0; m < u->n_members && err == hipSuccess; ++m)
    if (rows[m] > 0)
      err = hipMemcpyAsync(u->d_rows[m], u->h_rows[m], (size_t)rows[m] * GWL_NPARAM * 8,
                           hipMemcpyHostToDevice, u->stream);
  if (err != hipSuccess) {
    (void)gwl::hip_fail(err, "upload");
    return fail(s, u, GWL_E_HIP, "gwl_server upload");
  }
  int rc = GWL_OK;
  if (u->kind == UNIT_EVENT) {
    rc = gwl_eval(u->ev, u->d_rows[0], rows[0], GWL_NPARAM, u->precision, u->d_out[0], nullptr,
                  u->stream);
  } else if (u->kind == UNIT_ROQ) {
    rc = gwl_roq_eval(u->roq, u->d_rows[0], rows[0], GWL_NPARAM, u->mode, u->d_out[0], u->stream);
  } else if (u->kind == UNIT_RELBIN) {
    rc = gwl_relbin_eval(u->rb, u->d_rows[0], rows[0], GWL_NPARAM, u->precision, u->d_out[0],
                         nullptr, u->stream);
  } else if (u->kind == UNIT_CALIB) {
    rc = gwl_calib_eval(u->cal, u->d_rows[0], rows[0], GWL_NPARAM, u->mode, u->d_out[0], nullptr,
                        u->stream);
  } else {
    rc = gwl_multi_set_rows(u->multi, rows.data(), u->stream);
    if (rc == GWL_OK) rc = gwl_multi_launch(u->multi, u->stream);
  }
  if (rc != GWL_OK) return fail(s, u, rc, "gwl_server evaluation");
  for (int m = 0; m < u->n_members && err == hipSuccess; ++m)
    if (rows[m] > 0)
      err = hipMemcpyAsync(u->h_out[m], u->d_out[m], (size_t)rows[m] * 8, hipMemcpyDeviceToHost,
                           u->stream);
  if (err == hipSuccess) err = hipEventRecord(u->done, u->stream);
  if (err != hipSuccess) {
    (void)gwl::hip_fail(err, "download");
    return fail(s, u, GWL_E_HIP, "gwl_server download");
  }
}

void complete(gwl_server* s, Unit* u) {
  for (const Req& r : u->inflight)
    std::memcpy(r.out, u->h_out[r.member] + r.off, (size_t)r.n * 8);
  finish(s, u, GWL_OK);
}

void loop(gwl_server* s) {
  (void)hipSetDevice(s->ctx->device);
  std::vector<std::vector<Req>> take;
  std::vector<Unit*> units;
  std::vector<int64_t> taken;
  for (;;) {
    bool launched = false, finished = false;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      units = s->units;
      bool busy = false;
      for (Unit* u : units) busy = busy || u->busy;
      if (s->pending.empty() && !busy) {
        if (s->stop) return;
        s->cv_work.wait(lk, [&] { return s->stop || !s->pending.empty(); });
        if (s->pending.empty()) return;
        units = s->units;
      }
      take.resize(units.size());
      taken.assign(s->ids.size(), 0);
      std::deque<Req> keep;
      while (!s->pending.empty()) {
        Req r = s->pending.front();
        s->pending.pop_front();
        Unit* u = units[r.unit];
        int64_t& have = taken[r.id];  // rows already taken for this id in this pass
        if (u->busy || (have > 0 && have + r.n > s->max_rows)) {
          keep.push_back(r);
          continue;
        }
        have += r.n;
        take[r.unit].push_back(r);
      }
      s->pending.swap(keep);
    }
    for (size_t k = 0; k < units.size(); ++k) {
      if (take[k].empty()) continue;
      launch(s, units[k], take[k]);
      launched = true;
    }
    for (Unit* u : units) {
      if (!u->busy) continue;
      const hipError_t q = hipEventQuery(u->done);
      if (q == hipSuccess) {
        complete(s, u);
        finished = true;
      } else if (q != hipErrorNotReady) {
        (void)gwl::hip_fail(q, "hipEventQuery");
        fail(s, u, GWL_E_HIP, "gwl_server completion");
        finished = true;
      }
    }
    if (!launched && !finished) {
      std::unique_lock<std::mutex> lk(s->mu);
      if (s->pending.empty()) s->cv_work.wait_for(lk, std::chrono::microseconds(20));
    }
  }
}

int new_unit(gwl_server* s, Unit* u, int members) {
  u->n_members = members;
  GWL_HIP_CHECK(hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking));
  GWL_HIP_CHECK(hipEventCreateWithFlags(&u->done, hipEventDisableTiming));
  u->h_rows.assign(members, nullptr);
  u->h_out.assign(members, nullptr);
  u->d_rows.assign(members, nullptr);
  u->d_out.assign(members, nullptr);
  for (int m = 0; m < members; ++m) {
    GWL_HIP_CHECK(hipHostMalloc((void**)&u->h_rows[m], (size_t)s->max_rows * GWL_NPARAM * 8));
    GWL_HIP_CHECK(hipHostMalloc((void**)&u->h_out[m], (size_t)s->max_rows * 8));
    GWL_HIP_CHECK(hipMalloc((void**)&u->d_rows[m], (size_t)s->max_rows * GWL_NPARAM * 8));
    GWL_HIP_CHECK(hipMalloc((void**)&u->d_out[m], (size_t)s->max_rows * 8));
  }
  s->device_bytes += (int64_t)members * s->max_rows * (GWL_NPARAM + 1) * 8;
  return GWL_OK;
}

void free_unit(Unit* u) {
  if (u->multi) gwl_multi_destroy(u->multi);
  for (double* p : u->h_rows) if (p) (void)hipHostFree(p);
  for (double* p : u->h_out) if (p) (void)hipHostFree(p);
  for (double* p : u->d_rows) if (p) (void)hipFree(p);
  for (double* p : u->d_out) if (p) (void)hipFree(p);
  if (u->done) (void)hipEventDestroy(u->done);
  if (u->stream) (void)hipStreamDestroy(u->stream);
  delete u;
}

int add_unit(gwl_server* s, Unit* u, int members, int32_t* first_id) {
  std::lock_guard<std::mutex> lk(s->mu);
  const int idx = (int)s->units.size();
  s->units.push_back(u);
  *first_id = (int32_t)s->ids.size();
  for (int m = 0; m < members; ++m) s->ids.emplace_back(idx, m);
  return GWL_OK;
}

// ---------------------------------------------------------------- socket front-end
// Wire format (native byte order, both directions a 16-byte header):
//   request  {u32 magic GWL_WIRE_REQUEST, i32 id, i64 n} + n x 16 float64 rows
//   reply    {u32 magic GWL_WIRE_REPLY, i32 status, i64 n} + n float64 values
//            (status 0) or n bytes of error text (status != 0)
struct WireHeader {
  uint32_t magic;
  int32_t word;
  int64_t n;
};
static_assert(sizeof(WireHeader) == 16, "wire header is 16 bytes");

bool read_full(int fd, void* buf, size_t len) {
  char* p = static_cast<char*>(buf);
  while (len > 0) {
    const ssize_t k = recv(fd, p, len, 0);
    if (k > 0) { p += k; len -= (size_t)k; continue; }
    if (k < 0 && errno == EINTR) continue;
    return false;
  }
  return true;
}

bool write_full(int fd, const void* buf, size_t len) {
  const char* p = static_cast<const char*>(buf);
  while (len > 0) {
    const ssize_t k = send(fd, p, len, MSG_NOSIGNAL);
    if (k > 0) { p += k; len -= (size_t)k; continue; }
    if (k < 0 && errno == EINTR) continue;
    return false;
  }
  return true;
}

bool reply_error(int fd, int status, const std::string& msg) {
  const WireHeader h{GWL_WIRE_REPLY, status, (int64_t)msg.size()};
  return write_full(fd, &h, sizeof h) && write_full(fd, msg.data(), msg.size());
}

// One connected client: its requests are served in order, each through the
// same queue as gwl_server_submit, so rows from many clients share launches.
void conn_loop(gwl_server* s, SockConn* c) {
  std::vector<double> rows, out;
  const int fd = c->fd;
  for (;;) {
    WireHeader h;
    if (!read_full(fd, &h, sizeof h)) break;
    if (h.magic != GWL_WIRE_REQUEST || h.n < 1 || h.n > s->max_rows) {
      reply_error(fd, GWL_E_INVALID, "gwl_server socket: bad request header (magic, or n outside "
                                     "1..max_rows)");
      break;  // the stream cannot be resynchronised
    }
    rows.resize((size_t)h.n * GWL_NPARAM);
    out.resize((size_t)h.n);
    if (!read_full(fd, rows.data(), rows.size() * 8)) break;
    int64_t ticket = 0;
    int rc = gwl_server_submit(s, h.word, rows.data(), h.n, out.data(), &ticket);
    if (rc == GWL_OK) rc = gwl_server_wait(s, ticket, nullptr);
    if (rc != GWL_OK) {
      if (!reply_error(fd, rc, gwl_last_error())) break;
      continue;
    }
    const WireHeader r{GWL_WIRE_REPLY, 0, h.n};
    if (!write_full(fd, &r, sizeof r) || !write_full(fd, out.data(), out.size() * 8)) break;
  }
  std::lock_guard<std::mutex> lk(s->conn_mu);
  close(c->fd);
  c->fd = -1;
}

void accept_loop(gwl_server* s) {
  for (;;) {
    const int fd = accept4(s->listen_fd, nullptr, nullptr, SOCK_CLOEXEC);
    if (fd < 0) {
      if (errno == EINTR || errno == ECONNABORTED) continue;
      return;  // the listening socket was shut down
    }
    std::lock_guard<std::mutex> lk(s->conn_mu);
    if (s->closing) {
      close(fd);
      return;
    }
    SockConn* c = new SockConn;
    c->fd = fd;
    s->conns.push_back(c);
    c->th = std::thread(conn_loop, s, c);
  }
}

// Stop accepting, end every connection (a request in flight is finished by
// the scheduler first), remove the socket file.
void close_socket(gwl_server* s) {
  if (s->listen_fd < 0) return;
  {
    std::lock_guard<std::mutex> lk(s->conn_mu);
    s->closing = true;
  }
  shutdown(s->listen_fd, SHUT_RDWR);
  if (s->accept_th.joinable()) s->accept_th.join();
  close(s->listen_fd);
  s->listen_fd = -1;
  unlink(s->sock_path.c_str());
  std::vector<SockConn*> conns;
  {
    std::lock_guard<std::mutex> lk(s->conn_mu);
    conns.swap(s->conns);
    for (SockConn* c : conns)
      if (c->fd >= 0) shutdown(c->fd, SHUT_RDWR);
  }
  for (SockConn* c : conns) {
    if (c->th.joinable()) c->th.join();
    delete c;
  }
}

}  // namespace

using gwl::set_error;

extern "C" {

int gwl_server_create(gwl_ctx* ctx, int64_t max_rows, gwl_server** out) {
  if (!ctx || max_rows < 1 || !out) {
    set_error("gwl_server_create: invalid argument");
    return GWL_E_INVALID;
  }
  gwl_server* s = new (std::nothrow) gwl_server;
  if (!s) { set_error("out of host memory"); return GWL_E_NOMEM; }
  s->ctx = ctx;
  s->max_rows = max_rows;
  s->th = std::thread(loop, s);
  *out = s;
  return GWL_OK;
}

int gwl_server_destroy(gwl_server* s) {
  if (!s) return GWL_OK;
  close_socket(s);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->th.joinable()) s->th.join();
  (void)hipSetDevice(s->ctx->device);
  for (Unit* u : s->units) {
    if (u->busy) (void)hipEventSynchronize(u->done);
    free_unit(u);
  }
  delete s;
  return GWL_OK;
}

int gwl_server_add_event(gwl_server* s, gwl_event* ev, int32_t precision, int32_t* id) {
  if (!s || !ev || !id || (precision != GWL_FP32 && precision != GWL_FP64)) {
    set_error("gwl_server_add_event: invalid argument");
    return GWL_E_INVALID;
  }
  GWL_HIP_CHECK(hipSetDevice(s->ctx->device));
  if (int rc = gwl_event_reserve(ev, s->max_rows)) return rc;
  Unit* u = new Unit;
  u->kind = UNIT_EVENT;
  u->ev = ev;
  u->precision = precision;
  if (int rc = new_unit(s, u, 1)) {
    free_unit(u);
    return rc;
  }
  return add_unit(s, u, 1, id);
}

int gwl_server_add_roq(gwl_server* s, gwl_roq* roq, int32_t mode, int32_t* id) {
  if (!s || !roq || !id || mode < 0 || mode > 2) {
    set_error("gwl_server_add_roq: invalid argument");
    return GWL_E_INVALID;
  }
  GWL_HIP_CHECK(hipSetDevice(s->ctx->device));
  if (int rc = gwl_roq_reserve(roq, s->max_rows)) return rc;
  Unit* u = new Unit;
  u->kind = UNIT_ROQ;
  u->roq = roq;
  u->mode = mode;
  if (int rc = new_unit(s, u, 1)) {
    free_unit(u);
    return rc;
  }
  return add_unit(s, u, 1, id);
}

int gwl_server_add_relbin(gwl_server* s, gwl_relbin* rb, int32_t precision, int32_t* id) {
  if (!s || !rb || !id || (precision != GWL_FP32 && precision != GWL_FP64)) {
    set_error("gwl_server_add_relbin: invalid argument");
    return GWL_E_INVALID;
  }
  GWL_HIP_CHECK(hipSetDevice(s->ctx->device));
  Unit* u = new Unit;
  u->kind = UNIT_RELBIN;
  u->rb = rb;
  u->precision = precision;
  if (int rc = new_unit(s, u, 1)) {
    free_unit(u);
    return rc;
  }
  return add_unit(s, u, 1, id);
}

int gwl_server_add_calib(gwl_server* s, gwl_calib* cal, int32_t mode, int32_t* id) {
  if (!s || !cal || !id || mode < 0 || mode > 2) {
    set_error("gwl_server_add_calib: invalid argument");
    return GWL_E_INVALID;
  }
  GWL_HIP_CHECK(hipSetDevice(s->ctx->device));
  if (int rc = gwl_calib_reserve(cal, s->max_rows)) return rc;
  Unit* u = new Unit;
  u->kind = UNIT_CALIB;
  u->cal = cal;
  u->mode = mode;
  if (int rc = new_unit(s, u, 1)) {
    free_unit(u);
    return rc;
  }
  return add_unit(s, u, 1, id);
}

int gwl_server_add_group(gwl_server* s, int32_t n, gwl_event* const* events, int32_t* first_id) {
  if (!s || n < 1 || !events || !first_id) {
    set_error("gwl_server_add_group: invalid argument");
    return GWL_E_INVALID;
  }
  GWL_HIP_CHECK(hipSetDevice(s->ctx->device));
  Unit* u = new Unit;
  u->kind = UNIT_GROUP;
  if (int rc = new_unit(s, u, n)) {
    free_unit(u);
    return rc;
  }
  std::vector<int64_t> maxr(n, s->max_rows);
  std::vector<const double*> params(u->d_rows.begin(), u->d_rows.end());
  int rc = gwl_multi_create(s->ctx, n, events, params.data(), u->d_out.data(), maxr.data(),
                            &u->multi);
  if (rc != GWL_OK) {
    free_unit(u);
    return rc;
  }
  int64_t b = (int64_t)u->multi->dev_bytes;
  for (int e = 0; e < n; ++e) {
    const int64_t L = events[e]->dev.L;
    b += s->max_rows * 16 * (L * 16 + 8);
    if (u->multi->wave_path) b += (int64_t)gwl::roww_bytes(s->max_rows);
  }
  s->device_bytes += b;
  return add_unit(s, u, n, first_id);
}

int gwl_server_device_bytes(gwl_server* s, int64_t* bytes) {
  if (!s || !bytes) { set_error("gwl_server_device_bytes: invalid argument"); return GWL_E_INVALID; }
  std::lock_guard<std::mutex> lk(s->mu);
  *bytes = s->device_bytes;
  return GWL_OK;
}

int gwl_server_submit(gwl_server* s, int32_t id, const double* rows, int64_t n, double* out,
                      int64_t* ticket) {
  if (!s || !rows || !out || !ticket || n < 1) {
    set_error("gwl_server_submit: invalid argument");
    return GWL_E_INVALID;
  }
  Req r;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (id < 0 || id >= (int32_t)s->ids.size()) {
      set_error("gwl_server_submit: unknown id");
      return GWL_E_INVALID;
    }
    if (n > s->max_rows) {
      set_error("gwl_server_submit: more rows than the server's max_rows");
      return GWL_E_INVALID;
    }
    if (s->stop) {
      set_error("gwl_server_submit: server is stopping");
      return GWL_E_INVALID;
    }
    r.id = id;
    r.unit = s->ids[id].first;
    r.member = s->ids[id].second;
    r.rows = rows;
    r.n = n;
    r.out = out;
    r.off = 0;
    r.ticket = s->next_ticket++;
    r.t0 = Clock::now();
    s->tickets[r.ticket] = Ticket();
    s->pending.push_back(r);
  }
  s->cv_work.notify_one();
  *ticket = r.ticket;
  return GWL_OK;
}

int gwl_server_wait(gwl_server* s, int64_t ticket, double* latency_s) {
  if (!s) { set_error("gwl_server_wait: NULL server"); return GWL_E_INVALID; }
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->tickets.find(ticket);
  if (it == s->tickets.end()) {
    set_error("gwl_server_wait: unknown ticket");
    return GWL_E_INVALID;
  }
  s->cv_done.wait(lk, [&] { return s->tickets[ticket].done; });
  const Ticket t = s->tickets[ticket];
  s->tickets.erase(ticket);
  if (latency_s) *latency_s = t.latency;
  if (t.status != GWL_OK) set_error(s->last_error);
  return t.status;
}

int gwl_server_listen(gwl_server* s, const char* path) {
  sockaddr_un addr;
  std::memset(&addr, 0, sizeof addr);
  if (!s || !path || !*path || std::strlen(path) >= sizeof(addr.sun_path)) {
    set_error("gwl_server_listen: invalid argument (path empty or longer than a socket path)");
    return GWL_E_INVALID;
  }
  if (s->listen_fd >= 0) {
    set_error("gwl_server_listen: already listening");
    return GWL_E_INVALID;
  }
  addr.sun_family = AF_UNIX;
  std::strcpy(addr.sun_path, path);
  const int fd = socket(AF_UNIX, SOCK_STREAM | SOCK_CLOEXEC, 0);
  if (fd < 0) {
    set_error(std::string("gwl_server_listen: socket: ") + std::strerror(errno));
    return GWL_E_INVALID;
  }
  if (bind(fd, (const sockaddr*)&addr, sizeof addr) != 0 || chmod(path, 0600) != 0 ||
      listen(fd, 256) != 0) {
    set_error(std::string("gwl_server_listen: ") + path + ": " + std::strerror(errno));
    close(fd);
    return GWL_E_INVALID;
  }
  s->listen_fd = fd;
  s->sock_path = path;
  s->closing = false;
  s->accept_th = std::thread(accept_loop, s);
  return GWL_OK;
}

int gwl_server_stats(gwl_server* s, int64_t* batches, int64_t* rows) {
  if (!s || !batches || !rows) { set_error("gwl_server_stats: invalid argument"); return GWL_E_INVALID; }
  std::lock_guard<std::mutex> lk(s->mu);
  *batches = s->batches;
  *rows = s->rows_done;
  return GWL_OK;
}

}  // extern "C"
