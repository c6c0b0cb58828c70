// driver.cpp -- the simulation driver of the reference (example/example.cpp)
// on the MI355X hot path: derives the deck, builds the initial state on the
// host through the deck's initializers, uploads it, runs the step loop through
// the C-ABI and writes energies.txt, then calls the deck's finalizer.
//
// Usage: cbnpic [--stages] [--chunk N] [--dumps] [--dump-initial FILE]
//   default   the fused cpic_step, energies of every step kept on the device
//             and written after each chunk of N steps (default 200)
//   --stages  one C-ABI call per stage and step, in the reference's order
//             (example/example.cpp:216-279), energies read back every step
//   --dumps   the reference's per-step partloc and ex1d files as well: the
//             steps are queued without waiting, each followed by a snapshot
//             into one of two slots, and a writer thread formats step k
//             while the following steps run
//   --dump-initial FILE
//             write the derived constants and the initial particles to FILE
//             and exit without touching the GPU (checks of the deck layer)
//   --dump-constants FILE
//             the same without the particles
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cabanapic_amd.h"
#include "deck.h"

Input_Deck deck;

namespace {

void die(const char* what, int code) {
  std::fprintf(stderr, "cbnpic: %s failed (%d): %s\n", what, code, cpic_last_error());
  std::exit(1);
}

#define CPIC_CALL(expr)              \
  do {                               \
    const int rc_ = (expr);          \
    if (rc_ != CPIC_OK) die(#expr, rc_); \
  } while (0)

// energies.txt in the reference's format (src/fields.h:741-762): the file is
// truncated by the first step and appended afterwards; each line is also
// printed as "%d %f %e %e".  Built with ES_FIELD_SOLVER (the electrostatic
// 1D solver) a line has no B energy, as in the reference.
void dump_energy_line(int step, real_t time, real_t e, real_t b) {
  std::ofstream f;
  if (step == 1)
    f.open("energies.txt", std::ofstream::out | std::ofstream::trunc);
  else
    f.open("energies.txt", std::ios::app);
#ifdef ES_FIELD_SOLVER
  (void)b;
  f << step << " " << time << " " << e << std::endl;
  std::printf("%d %f %e\n", step, (double)time, (double)e);
#else
  f << step << " " << time << " " << e << " " << b << std::endl;
  std::printf("%d %f %e %e\n", step, (double)time, (double)e, (double)b);
#endif
}

// Workers that format parts of one dump line together: run(f) calls f(i)
// for i = 0..size()-1, f(0) on the caller, and returns when all are done.
class LinePool {
 public:
  explicit LinePool(unsigned n) {
    for (unsigned i = 1; i < n; ++i) th_.emplace_back([this, i] { loop(i); });
  }
  ~LinePool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  unsigned size() const { return (unsigned)th_.size() + 1; }
  void run(const std::function<void(unsigned)>& f) {
    {
      std::lock_guard<std::mutex> lk(m_);
      job_ = &f;
      ++gen_;
      left_ = (unsigned)th_.size();
    }
    cv_.notify_all();
    f(0);
    std::unique_lock<std::mutex> lk(m_);
    done_.wait(lk, [&] { return left_ == 0; });
  }

 private:
  void loop(unsigned i) {
    unsigned long seen = 0;
    for (;;) {
      const std::function<void(unsigned)>* f;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        f = job_;
      }
      (*f)(i);
      std::lock_guard<std::mutex> lk(m_);
      if (--left_ == 0) done_.notify_all();
    }
  }
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  const std::function<void(unsigned)>* job_ = nullptr;
  unsigned long gen_ = 0;
  unsigned left_ = 0;
  bool stop_ = false;
};

// dump_particles, src/helpers.h:31-70: the physical x position and ux of
// every particle, "%e  %e " each, then a newline.  With a pool, the parts of
// the line are formatted in parallel and written in order.
void write_particles(FILE* fp, size_t n, const real_t* px, const real_t* ux,
                     const int32_t* cell, real_t xmin, real_t dx, size_t nx, size_t ny,
                     size_t ng, LinePool* pool = nullptr) {
  const unsigned parts = pool && n >= 2048 ? pool->size() : 1;
  std::vector<std::string> text(parts);
  auto format = [&](unsigned t) {
    const size_t lo = n * t / parts, hi = n * (t + 1) / parts;
    std::string& out = text[t];
    out.reserve((hi - lo) * 30);
    char buf[80];
    for (size_t p = lo; p < hi; ++p) {
      size_t ix, iy, iz;
      const int ii = cell[p];
      RANK_TO_INDEX(ii, ix, iy, iz, nx + 2 * ng, ny + 2 * ng);
      (void)iy;
      (void)iz;
      const real_t x = xmin + (ix - 1 + (px[p] + 1.0) * 0.5) * dx;
      const real_t v = ux[p];
      const int k = std::snprintf(buf, sizeof buf, "%e  %e ", (double)x, (double)v);
      out.append(buf, (size_t)k);
    }
  };
  if (parts > 1)
    pool->run(format);
  else
    format(0);
  for (const auto& t : text) std::fwrite(t.data(), 1, t.size(), fp);
  std::fprintf(fp, "\n");
}

// Field_Solver::dump_fields, src/fields.h:318-350: ex along the line y = z = 1;
// ex_line[i - 1] is ex at voxel (i, 1, 1)
void write_fields(FILE* fp, const real_t* ex_line, real_t xmin, real_t dx, size_t nx) {
  for (size_t i = 1; i < nx + 1; i++) {
    const real_t x = xmin + (i - 0.5) * dx;
    std::fprintf(fp, "%e %e\n", (double)x, (double)ex_line[i - 1]);
  }
  std::fprintf(fp, "\n\n");
}

}  // namespace

int main(int argc, char** argv) {
  bool stages = false;
  long chunk = 200;
  const char* dump = nullptr;
  bool dump_particles = true;
  bool dumps = false;
  for (int i = 1; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--stages")) {
      stages = true;
    } else if (!std::strcmp(argv[i], "--dumps")) {
      dumps = true;
    } else if (!std::strcmp(argv[i], "--chunk") && i + 1 < argc) {
      chunk = std::atol(argv[++i]);
      if (chunk < 1) chunk = 1;
    } else if (!std::strcmp(argv[i], "--dump-initial") && i + 1 < argc) {
      dump = argv[++i];
    } else if (!std::strcmp(argv[i], "--dump-constants") && i + 1 < argc) {
      dump = argv[++i];
      dump_particles = false;
    } else {
      std::fprintf(stderr,
                   "usage: %s [--stages] [--chunk N] [--dumps] "
                   "[--dump-initial|--dump-constants FILE]\n",
                   argv[0]);
      return 2;
    }
  }
  std::printf("#Running on the MI355X hot path (real_t = %s)\n",
              sizeof(real_t) == 4 ? "float" : "double");

  deck.derive_params();
  deck.print_run_details();

  // driver constants with the reference's types (example/example.cpp:65-114)
  const int npc = (int)deck.nppc;
  const int nx = (int)deck.nx, ny = (int)deck.ny, nz = (int)deck.nz;
  const int num_ghosts = (int)deck.num_ghosts;
#ifdef ES_FIELD_SOLVER
  // example/example.cpp:69-74
  if (ny > 1 || nz > 1) {
    std::fprintf(stderr, "Error: ES Field solver supports 1D only.\n");
    return -1;
  }
#endif
  const size_t num_cells = deck.num_cells;
  real_t dxp = 2.f / (npc);
  const real_t dx = deck.dx, dy = deck.dy, dz = deck.dz;
  real_t dt = deck.dt;
  real_t c = deck.c;
  real_t Lx = deck.len_x, Ly = deck.len_y, Lz = deck.len_z;
  real_t v0 = deck.v0;
  real_t eps0 = deck.eps;
  real_t Npe = deck.Npe;
  size_t Ne = (size_t)deck.Ne;
  real_t qsp = deck.qsp;
  real_t me = deck.me;
  real_t qdt_2mc = qsp * dt / (2 * me * c);
  real_t cdt_dx = c * dt / dx;
  real_t cdt_dy = c * dt / dy;
  real_t cdt_dz = c * dt / dz;
  real_t dt_eps0 = dt / eps0;
  real_t frac = 1.0f;
  real_t we = (real_t)Npe / (real_t)Ne;
  const size_t num_particles = (size_t)deck.num_particles;
  std::printf("nppc %d nx %d ny %d nz %d \n", npc, nx, ny, nz);
  std::printf("ne %zu npe %e we %e qsp %e\n", Ne, (double)Npe, (double)we, (double)qsp);
  std::printf("c %e dt %e dx %e cdt_dx %e \n", (double)c, (double)dt, (double)dx, (double)cdt_dx);

  particle_list_t particles(dump && !dump_particles ? 0 : num_particles);
  if (!dump || dump_particles)
    deck.initialize_particles(particles, nx, ny, nz, num_ghosts, dxp, npc, we, v0);
  field_array_t fields(num_cells);
  deck.initialize_fields(fields, nx, ny, nz, num_ghosts, Lx, Ly, Lz, dx, dy, dz);
  const Boundary boundary = deck.BOUNDARY_TYPE;

  // (example/example.cpp:179-181)
  const real_t px = (nx > 1) ? frac * c * dt / dx : 0;
  const real_t py = (ny > 1) ? frac * c * dt / dy : 0;
  const real_t pz = (nz > 1) ? frac * c * dt / dz : 0;
  const int num_steps = deck.num_steps;
  std::printf("#num_step = %d\n#dt*wpe = %f\n#we = %f\n", num_steps, (double)dt, (double)we);

  if (dump) {
    // int64 np, int32 sizeof(real_t), 16 constants as double, then the
    // particle members x,y,z,ux,uy,uz,w (real_t) and the int32 cells
    std::FILE* f = std::fopen(dump, "wb");
    if (!f) return 1;
    const int64_t n = dump_particles ? (int64_t)num_particles : 0;
    const int32_t rs = (int32_t)sizeof(real_t);
    const double k[16] = {qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp, dx,  dy, dz,
                          dt,      px,     py,     pz,     dt_eps0, dxp, we, (double)num_steps};
    std::fwrite(&n, sizeof n, 1, f);
    std::fwrite(&rs, sizeof rs, 1, f);
    std::fwrite(k, sizeof k, 1, f);
    for (int m = PositionX; m <= Weight; ++m)
      std::fwrite(particles.data(m), sizeof(real_t), (size_t)n, f);
    std::fwrite(particles.cells(), sizeof(int32_t), (size_t)n, f);
    std::fclose(f);
    return 0;
  }

  if (boundary != Boundary::Periodic) {
    // the reference's field solver exits for anything but periodic
    // boundaries (src/fields.h:21-25,113-117)
    std::fprintf(stderr, "cbnpic: only periodic boundaries are supported\n");
    return 1;
  }

  // the reference's per-step dumps (example/example.cpp:59-60,130-132,
  // 273-277): partloc (particles) and ex1d (ex along x), with --dumps
  FILE* fptr = nullptr;
  FILE* fpfd = nullptr;
  std::unique_ptr<LinePool> pool;
  if (dumps) {
    pool.reset(new LinePool(std::max(1u, std::min(8u, std::thread::hardware_concurrency() / 2))));
    fptr = std::fopen("partloc", "w");
    fpfd = std::fopen("ex1d", "w");
    if (!fptr || !fpfd) {
      std::fprintf(stderr, "cbnpic: cannot open partloc / ex1d\n");
      return 1;
    }
  }
  // one step's dump from a snapshot (energies line included when asked)
  auto write_dump = [&](const cpic_snapshot& sn, int step, bool energies) {
    if (energies)
      dump_energy_line(step, step * dt, (real_t)sn.e_energy, (real_t)sn.b_energy);
    std::fprintf(fpfd, "#step=%d\n", step);
    write_fields(fpfd, (const real_t*)sn.ex_line, 0, dx, nx);
    std::fprintf(fptr, "#step=%d\n%e ", step, (double)(step * dt));
    write_particles(fptr, (size_t)sn.n, (const real_t*)sn.x, (const real_t*)sn.ux, sn.cell, 0,
                    dx, nx, ny, num_ghosts, pool.get());
  };

  cpic_config cfg{};
  cfg.real_size = (int32_t)sizeof(real_t);
  cfg.num_ghosts = num_ghosts;
  cfg.nx = nx;
  cfg.ny = ny;
  cfg.nz = nz;
  cfg.max_particles = (int64_t)num_particles;
  cpic_context* ctx = nullptr;
  CPIC_CALL(cpic_create(&cfg, &ctx));
#ifdef ES_FIELD_SOLVER
  // Field_Solver<ES_Field_Solver_1D>, example/example.cpp:27-29
  CPIC_CALL(cpic_set_field_solver(ctx, CPIC_FIELD_SOLVER_ES_1D));
#endif
  CPIC_CALL(cpic_upload_particles(ctx, (int64_t)num_particles, particles.data(PositionX),
                                  particles.data(PositionY), particles.data(PositionZ),
                                  particles.data(VelocityX), particles.data(VelocityY),
                                  particles.data(VelocityZ), particles.data(Weight),
                                  particles.cells()));
  CPIC_CALL(cpic_upload_fields(ctx, fields.data()));

  // the step-0 particle line of the dumps (example/example.cpp:130-132), taken
  // before the uncentering from a snapshot of the uploaded particles (upload
  // order is the order of the initial arrays); it also sets up both snapshot
  // slots before the steps
  if (dumps) {
    cpic_snapshot sn{};
    CPIC_CALL(cpic_snapshot_begin(ctx, 0));
    CPIC_CALL(cpic_snapshot_wait(ctx, 0, &sn));
    std::fprintf(fptr, "#step=0\n0 ");
    write_particles(fptr, (size_t)sn.n, (const real_t*)sn.x, (const real_t*)sn.ux, sn.cell, 0, dx,
                    nx, ny, num_ghosts, pool.get());
  }

  // example/example.cpp:204-213
  if (deck.perform_uncenter) {
    CPIC_CALL(cpic_load_interpolator_array(ctx));
    CPIC_CALL(cpic_uncenter_particles(ctx, qdt_2mc));
  }

  // Gauss's law in FP64 on the device: the charge-conserving deposit keeps
  // div E - rho / eps0 constant at every node; its change is printed at the end
  double gauss[2] = {0, 0};
  CPIC_CALL(cpic_gauss_residual(ctx, dx, dy, dz, eps0, qsp, CPIC_GAUSS_BASELINE, gauss, nullptr));

  const auto loop_start = std::chrono::steady_clock::now();
  if (stages) {
    for (int step = 1; step <= num_steps; step++) {
      CPIC_CALL(cpic_load_interpolator_array(ctx));
      CPIC_CALL(cpic_clear_accumulator_array(ctx));
      CPIC_CALL(cpic_push(ctx, qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp, CPIC_BOUNDARY_PERIODIC));
      CPIC_CALL(cpic_unload_accumulator_array(ctx, dx, dy, dz, dt));
      CPIC_CALL(cpic_advance_b(ctx, real_t(0.5) * px, real_t(0.5) * py, real_t(0.5) * pz));
      CPIC_CALL(cpic_advance_e(ctx, px, py, pz, dt_eps0));
      CPIC_CALL(cpic_advance_b(ctx, real_t(0.5) * px, real_t(0.5) * py, real_t(0.5) * pz));
      double e = 0, b = 0;
      CPIC_CALL(cpic_energies(ctx, &e, &b));
      dump_energy_line(step, step * dt, (real_t)e, (real_t)b);
      if (dumps) {
        cpic_snapshot sn{};
        CPIC_CALL(cpic_snapshot_begin(ctx, 0));
        CPIC_CALL(cpic_snapshot_wait(ctx, 0, &sn));
        write_dump(sn, step, false);
      }
    }
  } else {
    cpic_step_params sp{};
    sp.qdt_2mc = qdt_2mc;
    sp.cdt_dx = cdt_dx;
    sp.cdt_dy = cdt_dy;
    sp.cdt_dz = cdt_dz;
    sp.qsp = qsp;
    sp.dx = dx;
    sp.dy = dy;
    sp.dz = dz;
    sp.dt = dt;
    sp.px = px;
    sp.py = py;
    sp.pz = pz;
    sp.dt_eps0 = dt_eps0;
    sp.me = me;
    sp.c = c;
    sp.boundary = CPIC_BOUNDARY_PERIODIC;
    sp.energy_interval = 1;
    sp.compact_interval = 0;
    sp.kinetic_interval = 0;
    if (dumps) {
      // step k is queued, then its snapshot into slot k % 2; the writer
      // formats step k from that slot while steps k + 1, k + 2 run.  A slot
      // is begun again only after the writer is done with it.
      std::mutex m;
      std::condition_variable cv;
      int begun = 0, written = 0;
      bool failed = false;
      std::thread writer([&] {
        for (int step = 1; step <= num_steps; ++step) {
          {
            std::unique_lock<std::mutex> lk(m);
            cv.wait(lk, [&] { return begun >= step || failed; });
            if (failed) return;
          }
          cpic_snapshot sn{};
          const int rc = cpic_snapshot_wait(ctx, step % 2, &sn);
          if (rc == CPIC_OK) write_dump(sn, step, true);
          std::lock_guard<std::mutex> lk(m);
          if (rc != CPIC_OK) {
            std::fprintf(stderr, "cbnpic: cpic_snapshot_wait failed (%d): %s\n", rc,
                         cpic_last_error());
            failed = true;
          } else {
            written = step;
          }
          cv.notify_all();
          if (failed) return;
        }
      });
      int rc = CPIC_OK;
      const char* what = "";
      for (int step = 1; step <= num_steps && rc == CPIC_OK; ++step) {
        if ((rc = cpic_step_async(ctx, &sp, 1)) != CPIC_OK) {
          what = "cpic_step_async";
          break;
        }
        std::unique_lock<std::mutex> lk(m);
        cv.wait(lk, [&] { return written >= step - 2 || failed; });
        if (failed) break;
        if ((rc = cpic_snapshot_begin(ctx, step % 2)) != CPIC_OK) {
          what = "cpic_snapshot_begin";
          break;
        }
        begun = step;
        cv.notify_all();
      }
      {
        std::lock_guard<std::mutex> lk(m);
        if (rc != CPIC_OK) failed = true;
        cv.notify_all();
      }
      writer.join();
      if (rc != CPIC_OK) die(what, rc);
      if (failed) std::exit(1);
    } else {
      std::vector<double> hist;
      for (int first = 1; first <= num_steps; first += (int)chunk) {
        const int n = (int)std::min<long>(chunk, num_steps - first + 1);
        hist.assign(3 * (size_t)n, 0.0);
        CPIC_CALL(cpic_step(ctx, &sp, n, hist.data()));
        for (int k = 0; k < n; ++k) {
          const int step = first + k;
          dump_energy_line(step, step * dt, (real_t)hist[3 * k], (real_t)hist[3 * k + 1]);
        }
      }
    }
    CPIC_CALL(cpic_synchronize(ctx));
  }
  std::fprintf(stderr, "#step loop: %.6f s for %d steps\n",
               std::chrono::duration<double>(std::chrono::steady_clock::now() - loop_start)
                   .count(),
               num_steps);
  if (fptr) std::fclose(fptr);
  if (fpfd) std::fclose(fpfd);
  CPIC_CALL(cpic_gauss_residual(ctx, dx, dy, dz, eps0, qsp, CPIC_GAUSS_CHANGE, gauss, nullptr));
  std::fprintf(stderr, "#Gauss residual: max |change| %e, max |rho/eps0| %e (relative %e)\n",
               gauss[0], gauss[1], gauss[1] > 0 ? gauss[0] / gauss[1] : 0.0);
  int32_t flags = 0;
  CPIC_CALL(cpic_device_errors(ctx, &flags));
  CPIC_CALL(cpic_destroy(ctx));
  if (flags) {
    std::fprintf(stderr, "cbnpic: device error flags %d\n", flags);
    return 1;
  }

  // let the deck check or post-process the run (example/example.cpp:287)
  deck.finalize();
  return 0;
}
