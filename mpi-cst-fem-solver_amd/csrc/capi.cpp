// capi.cpp -- the extern "C" boundary (include/cstfem_hip.h): argument checks, context life
// cycle, host<->device copies and the host-side pieces of the path (traction loads, the
// generic CSR -> SELL conversion).  Every entry point converts exceptions into a status
// code plus a thread-local message.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <unordered_map>

#include "internal.hpp"

namespace cstfem_hip {

namespace {
thread_local std::string g_error;
}

void set_error(const std::string& m) { g_error = m; }

void hip_check(hipError_t e, const char* what, const char* file, int line) {
  if (e != hipSuccess)
    throw Error(CSTFEM_ERR_DEVICE, std::string("HIP error '") + hipGetErrorString(e) + "' in " +
                                       what + " (" + file + ":" + std::to_string(line) + ")");
}

unsigned char* Context::stage(size_t bytes) {
  if (bytes > h_stage_n) {
    if (h_stage) {
      HIP_CHECK(hipStreamSynchronize(stream));
      (void)hipHostFree(h_stage);
      h_stage = nullptr;
      h_stage_n = 0;
    }
    const size_t n = std::max<size_t>(bytes, 64 * 1024);
    HIP_CHECK(hipHostMalloc((void**)&h_stage, n, hipHostMallocDefault));
    h_stage_n = n;
  }
  return h_stage;
}

Context::~Context() {
  mg.join_prepared();  // its download runs on the auxiliary stream
  if (stream) (void)hipStreamSynchronize(stream);
  drop_batch_graph();
  if (h_stage) (void)hipHostFree(h_stage);
  if (h_scal) (void)hipHostFree(h_scal);
  if (aux) (void)hipStreamDestroy(aux);
  if (stream) (void)hipStreamDestroy(stream);
}

namespace {

template <class F>
int guarded(F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    g_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_error = "host allocation failed";
    return CSTFEM_ERR_ALLOC;
  } catch (const std::exception& e) {
    g_error = e.what();
    return CSTFEM_ERR_ARGUMENT;
  }
}

void need(bool ok, const std::string& msg) {
  if (!ok) throw Error(CSTFEM_ERR_ARGUMENT, msg);
}

void use_device(Context* c) { HIP_CHECK(hipSetDevice(c->device)); }

}  // namespace
}  // namespace cstfem_hip

using namespace cstfem_hip;

struct cstfem_ctx {
  Context c;
};

extern "C" {

const char* cstfem_last_error(void) { return cstfem_hip::g_error.c_str(); }

int cstfem_ctx_create(int device, cstfem_ctx** out) {
  return guarded([&] {
    need(out != nullptr, "cstfem_ctx_create: out is NULL");
    int count = 0;
    HIP_CHECK(hipGetDeviceCount(&count));
    need(device >= 0 && device < count, "cstfem_ctx_create: device " + std::to_string(device) +
                                            " not present (" + std::to_string(count) + " visible)");
    auto* x = new cstfem_ctx;
    Context& c = x->c;
    try {
      c.device = device;
      HIP_CHECK(hipSetDevice(device));
      HIP_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
      hipDeviceProp_t prop;
      HIP_CHECK(hipGetDeviceProperties(&prop, device));
      c.cu_count = prop.multiProcessorCount;
      HIP_CHECK(hipHostMalloc((void**)&c.h_scal, 2 * sizeof(PcgScalars), hipHostMallocDefault));
      c.scal.alloc(1);
      c.err_flag.alloc(4);
      noalloc_attach(c);
    } catch (...) {
      delete x;
      throw;
    }
    *out = x;
    return CSTFEM_OK;
  });
}

int cstfem_ctx_destroy(cstfem_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return CSTFEM_OK;
    (void)hipSetDevice(ctx->c.device);
    (void)hipStreamSynchronize(ctx->c.stream);
    delete ctx;
    return CSTFEM_OK;
  });
}

int cstfem_device_info(cstfem_ctx* ctx, char* name, int len, int* cu_count) {
  return guarded([&] {
    need(ctx != nullptr, "cstfem_device_info: NULL context");
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, ctx->c.device));
    if (name && len > 0) {
      std::snprintf(name, len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (cu_count) *cu_count = prop.multiProcessorCount;
    return CSTFEM_OK;
  });
}

int cstfem_upload_mesh(cstfem_ctx* ctx, int nn, const double* xy, int64_t ne, const int* tri) {
  return guarded([&] {
    need(ctx && xy && tri, "cstfem_upload_mesh: NULL argument");
    need(nn > 0 && ne > 0, "cstfem_upload_mesh: empty mesh");
    need((int64_t)2 * nn < INT32_MAX && 9 * ne < INT32_MAX, "cstfem_upload_mesh: mesh too large for 32-bit indices");
    for (int64_t k = 0; k < 3 * ne; ++k)
      need(tri[k] >= 0 && tri[k] < nn, "cstfem_upload_mesh: triangle " + std::to_string(k / 3) +
                                           " references node " + std::to_string(tri[k]) + " out of range");
    Context& c = ctx->c;
    use_device(&c);
    mg_drop_prepared(c);
    c.nn = nn;
    c.ne = ne;
    c.h_xy.assign(xy, xy + 2 * (size_t)nn);
    c.h_tri.assign(tri, tri + 3 * (size_t)ne);
    c.xy.alloc(2 * (size_t)nn);
    c.tri.alloc(3 * (size_t)ne);
    HIP_CHECK(hipMemcpyAsync(c.xy.get(), xy, sizeof(double) * 2 * nn, hipMemcpyHostToDevice, c.stream));
    HIP_CHECK(hipMemcpyAsync(c.tri.get(), tri, sizeof(int) * 3 * ne, hipMemcpyHostToDevice, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
    c.have_pattern = c.mesh_pattern = c.setup_done = c.vals32_valid = false;
    c.has_E_elem = c.has_nu_elem = false;
    c.h_segments.clear();
    c.batch_count = 0;
    c.fixed_valid = false;
    c.mpnn.built = false;
    c.mpnn.npairs = 0;
    c.ne_ptr.release();
    c.ne_list.release();
    c.trac_groups.clear();
    c.b.release();
    return CSTFEM_OK;
  });
}

int cstfem_set_material(cstfem_ctx* ctx, double E, double nu, int plane, double thickness,
                        const double* E_elem) {
  return guarded([&] {
    need(ctx != nullptr, "cstfem_set_material: NULL context");
    Context& c = ctx->c;
    use_device(&c);
    need(c.nn > 0, "cstfem_set_material: upload a mesh first");
    need(E > 0, "E must be positive");
    need(nu > -1.0 && nu < 0.5, "nu must be in (-1, 0.5)");
    need(thickness > 0, "thickness must be positive");
    need(plane == CSTFEM_PLANE_STRESS || plane == CSTFEM_PLANE_STRAIN,
         "plane must be plane_stress or plane_strain");
    c.E = E;
    c.nu = nu;
    c.thick = thickness;
    c.plane_strain = plane;
    c.D = constitutive_host(E, nu, plane);
    c.has_E_elem = E_elem != nullptr;
    c.has_nu_elem = false;
    if (E_elem) {
      for (int64_t e = 0; e < c.ne; ++e)
        need(E_elem[e] > 0, "element " + std::to_string(e) + ": E must be positive");
      c.E_elem.alloc(c.ne);
      HIP_CHECK(hipMemcpyAsync(c.E_elem.get(), E_elem, sizeof(double) * c.ne, hipMemcpyHostToDevice, c.stream));
      HIP_CHECK(hipStreamSynchronize(c.stream));
    }
    return CSTFEM_OK;
  });
}

int cstfem_set_constitutive(cstfem_ctx* ctx, const double* D, double thickness) {
  return guarded([&] {
    need(ctx && D, "cstfem_set_constitutive: NULL argument");
    need(thickness > 0, "thickness must be positive");
    Context& c = ctx->c;
    for (int i = 0; i < 9; ++i) {
      need(std::isfinite(D[i]), "cstfem_set_constitutive: D must be finite");
      c.D.d[i] = D[i];
    }
    c.thick = thickness;
    c.has_E_elem = c.has_nu_elem = false;
    return CSTFEM_OK;
  });
}

int cstfem_element_stiffness(cstfem_ctx* ctx, int64_t count, const double* xy, const double* D,
                             double thickness, double* Ke) {
  return guarded([&] {
    need(ctx && D && (count == 0 || (xy && Ke)), "cstfem_element_stiffness: NULL argument");
    need(count >= 0 && count <= (int64_t(1) << 40) / 42, "cstfem_element_stiffness: bad count");
    Context& c = ctx->c;
    use_device(&c);
    element_stiffness(c, count, xy, D, thickness, Ke);
    return CSTFEM_OK;
  });
}

int cstfem_build_pattern(cstfem_ctx* ctx, int64_t* nnz_out) {
  return guarded([&] {
    need(ctx != nullptr, "cstfem_build_pattern: NULL context");
    Context& c = ctx->c;
    use_device(&c);
    need(c.nn > 0, "cstfem_build_pattern: upload a mesh first");
    build_pattern(c);
    ensure_vectors(c);
    if (nnz_out) *nnz_out = 4 * c.nblk;
    return CSTFEM_OK;
  });
}

int cstfem_download_pattern(cstfem_ctx* ctx, int* row_ptr, int* cols) {
  return guarded([&] {
    need(ctx != nullptr, "cstfem_download_pattern: NULL context");
    Context& c = ctx->c;
    use_device(&c);
    need(c.have_pattern && c.mesh_pattern, "cstfem_download_pattern: no mesh pattern");
    std::vector<int> bp(c.nb + 1), bc(c.nblk);
    HIP_CHECK(hipMemcpyAsync(bp.data(), c.brow_ptr.get(), sizeof(int) * (c.nb + 1), hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipMemcpyAsync(bc.data(), c.bcol.get(), sizeof(int) * c.nblk, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
    // 2x expansion (assembly.cpp:21-39): rows 2n and 2n+1 share [2m, 2m+1 for m in adj(n)]
    if (row_ptr) {
      row_ptr[0] = 0;
      for (int nd = 0; nd < c.nb; ++nd) {
        const int len = 2 * (bp[nd + 1] - bp[nd]);
        row_ptr[2 * nd + 1] = row_ptr[2 * nd] + len;
        row_ptr[2 * nd + 2] = row_ptr[2 * nd + 1] + len;
      }
    }
    if (cols) {
      int64_t k = 0;
      for (int nd = 0; nd < c.nb; ++nd)
        for (int comp = 0; comp < 2; ++comp)
          for (int q = bp[nd]; q < bp[nd + 1]; ++q) {
            cols[k++] = 2 * bc[q];
            cols[k++] = 2 * bc[q] + 1;
          }
    }
    return CSTFEM_OK;
  });
}

int cstfem_assemble(cstfem_ctx* ctx) {
  return guarded([&] {
    need(ctx != nullptr, "cstfem_assemble: NULL context");
    use_device(&ctx->c);
    assemble_values(ctx->c);
    HIP_CHECK(hipStreamSynchronize(ctx->c.stream));
    return CSTFEM_OK;
  });
}

int cstfem_download_values(cstfem_ctx* ctx, double* vals) {
  return guarded([&] {
    need(ctx && vals, "cstfem_download_values: NULL argument");
    use_device(&ctx->c);
    download_values(ctx->c, vals);
    return CSTFEM_OK;
  });
}

int cstfem_upload_csr(cstfem_ctx* ctx, int n, const int* row_ptr, const int* cols,
                      const double* vals) {
  return guarded([&] {
    need(ctx && row_ptr && (n == 0 || (cols && vals)), "cstfem_upload_csr: NULL argument");
    need(n > 0, "cstfem_upload_csr: empty matrix");
    Context& c = ctx->c;
    use_device(&c);
    const int nb = (n + 1) / 2;
    // block rows: union of the block columns of rows 2I and 2I+1, ascending
    std::vector<int> bp(nb + 1, 0), bc;
    std::vector<int> bdiag(nb, -1);
    std::vector<int> tmp;
    for (int I = 0; I < nb; ++I) {
      tmp.clear();
      for (int rr = 2 * I; rr < std::min(2 * I + 2, n); ++rr) {
        need(row_ptr[rr + 1] >= row_ptr[rr], "cstfem_upload_csr: row_ptr not monotone");
        for (int k = row_ptr[rr]; k < row_ptr[rr + 1]; ++k) {
          need(cols[k] >= 0 && cols[k] < n, "cstfem_upload_csr: column out of range");
          if (k > row_ptr[rr]) need(cols[k] > cols[k - 1], "cstfem_upload_csr: columns not sorted");
          tmp.push_back(cols[k] / 2);
        }
      }
      std::sort(tmp.begin(), tmp.end());
      tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
      bp[I + 1] = bp[I] + (int)tmp.size();
      for (size_t q = 0; q < tmp.size(); ++q) {
        if (tmp[q] == I) bdiag[I] = bp[I] + (int)q;
        bc.push_back(tmp[q]);
      }
    }
    const int64_t nblk = bc.size();
    // structural symmetry of the block pattern: Dirichlet elimination then visits only the
    // rows named in the constrained nodes' rows, otherwise every row
    bool symmetric = true;
    int max_row = 0;
    for (int I = 0; I < nb && symmetric; ++I) {
      for (int q = bp[I]; q < bp[I + 1]; ++q) {
        const int J = bc[q];
        if (!std::binary_search(bc.begin() + bp[J], bc.begin() + bp[J + 1], I)) {
          symmetric = false;
          break;
        }
      }
    }
    for (int I = 0; I < nb; ++I) max_row = std::max(max_row, bp[I + 1] - bp[I]);
    // SELL-64 layout
    const int nslices = (nb + kSlice - 1) / kSlice;
    std::vector<int64_t> off(nslices + 1, 0);
    for (int s = 0; s < nslices; ++s) {
      int w = 0;
      for (int l = 0; l < kSlice && s * kSlice + l < nb; ++l) {
        const int r = s * kSlice + l;
        w = std::max(w, bp[r + 1] - bp[r]);
      }
      off[s + 1] = off[s] + (int64_t)w * kSlice;
    }
    const int64_t sb = off[nslices];
    std::vector<int> scol(sb);
    std::vector<double> sval(4 * sb, 0.0);
    std::vector<int64_t> spos(std::max<int64_t>(nblk, 1));
    for (int s = 0; s < nslices; ++s) {
      const int w = (int)((off[s + 1] - off[s]) / kSlice);
      for (int l = 0; l < kSlice; ++l) {
        const int r = s * kSlice + l;
        const int d = r < nb ? bp[r + 1] - bp[r] : 0;
        for (int k = 0; k < w; ++k) {
          const int64_t slot = off[s] + (int64_t)k * kSlice + l;
          if (k < d) {
            scol[slot] = bc[bp[r] + k];
            spos[bp[r] + k] = slot;
          } else {
            scol[slot] = r < nb ? r : 0;
          }
        }
      }
    }
    c.h_generic_map.assign(row_ptr[n], 0);
    for (int rr = 0; rr < n; ++rr) {
      const int I = rr / 2, ci = rr % 2;
      int q = bp[I];
      for (int k = row_ptr[rr]; k < row_ptr[rr + 1]; ++k) {
        const int J = cols[k] / 2, cj = cols[k] % 2;
        while (bc[q] != J) ++q;
        const int64_t vi = sell_vidx(spos[q], 2 * ci + cj);
        sval[vi] = vals[k];
        c.h_generic_map[k] = vi;
      }
    }
    c.nb = nb;
    c.n = n;
    c.nblk = nblk;
    c.nslices = nslices;
    c.sell_blocks = sb;
    c.h_sell_off = off;
    mg_drop_prepared(c);  // a level-0 aggregation under way reads the old pattern
    c.brow_ptr.alloc(nb + 1);
    c.bcol.alloc(std::max<int64_t>(nblk, 1));
    c.bdiag.alloc(nb);
    c.sell_off.alloc(nslices + 1);
    c.sell_col.alloc(std::max<int64_t>(sb, 1));
    c.sell_val.alloc(std::max<int64_t>(4 * sb, 1));
    c.sell_pos.alloc(spos.size());
    HIP_CHECK(hipMemcpyAsync(c.brow_ptr.get(), bp.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, c.stream));
    if (nblk)
      HIP_CHECK(hipMemcpyAsync(c.bcol.get(), bc.data(), sizeof(int) * nblk, hipMemcpyHostToDevice, c.stream));
    HIP_CHECK(hipMemcpyAsync(c.bdiag.get(), bdiag.data(), sizeof(int) * nb, hipMemcpyHostToDevice, c.stream));
    HIP_CHECK(hipMemcpyAsync(c.sell_off.get(), off.data(), sizeof(int64_t) * (nslices + 1), hipMemcpyHostToDevice, c.stream));
    if (sb) {
      HIP_CHECK(hipMemcpyAsync(c.sell_col.get(), scol.data(), sizeof(int) * sb, hipMemcpyHostToDevice, c.stream));
      HIP_CHECK(hipMemcpyAsync(c.sell_val.get(), sval.data(), sizeof(double) * 4 * sb, hipMemcpyHostToDevice, c.stream));
    }
    HIP_CHECK(hipMemcpyAsync(c.sell_pos.get(), spos.data(), sizeof(int64_t) * spos.size(), hipMemcpyHostToDevice, c.stream));
    std::vector<int> row0(nslices), nrow(nslices);
    for (int s = 0; s < nslices; ++s) {
      row0[s] = s * kSlice;
      nrow[s] = std::min(kSlice, nb - s * kSlice);
    }
    c.sell_row0.alloc(nslices + 1);
    c.sell_nrow.alloc(nslices + 1);
    HIP_CHECK(hipMemcpyAsync(c.sell_row0.get(), row0.data(), sizeof(int) * nslices, hipMemcpyHostToDevice, c.stream));
    HIP_CHECK(hipMemcpyAsync(c.sell_nrow.get(), nrow.data(), sizeof(int) * nslices, hipMemcpyHostToDevice, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
    c.h_segments.clear();
    c.h_seg_slice = {0, nslices};
    c.have_pattern = true;
    c.mesh_pattern = false;
    c.generic = true;
    c.pattern_symmetric = symmetric;
    c.max_row_blocks = max_row;
    c.fixed_valid = false;
    c.mpnn.built = false;
    c.mpnn.npairs = 0;
    c.setup_done = false;
    c.vals32_valid = false;
    c.b.release();
    ensure_vectors(c);
    return CSTFEM_OK;
  });
}

int cstfem_set_rhs(cstfem_ctx* ctx, const double* b, int n) {
  return guarded([&] {
    need(ctx != nullptr, "cstfem_set_rhs: NULL argument");
    Context& c = ctx->c;
    use_device(&c);
    need(c.have_pattern, "cstfem_set_rhs: no matrix on the device");
    need(n == c.n, "cstfem_set_rhs: size " + std::to_string(n) + " != " + std::to_string(c.n));
    ensure_vectors(c);
    HIP_CHECK(hipMemsetAsync(c.b.get(), 0, sizeof(double) * c.b.n, c.stream));
    // b == NULL: the zero vector the reference starts its load assembly from, set on the
    // device instead of copied from the host, and nothing to wait for (the library's later
    // work follows it on the stream); a host vector is waited for before it is handed back
    if (b) {
      HIP_CHECK(hipMemcpyAsync(c.b.get(), b, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
      HIP_CHECK(hipStreamSynchronize(c.stream));
    }
    HIP_CHECK(hipGetLastError());
    return CSTFEM_OK;
  });
}

int cstfem_download_rhs(cstfem_ctx* ctx, double* b, int n) {
  return guarded([&] {
    need(ctx && b, "cstfem_download_rhs: NULL argument");
    Context& c = ctx->c;
    use_device(&c);
    need(c.b.n > 0 && n == c.n, "cstfem_download_rhs: size mismatch or no RHS");
    HIP_CHECK(hipMemcpyAsync(b, c.b.get(), sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
    return CSTFEM_OK;
  });
}

int cstfem_add_body_force(cstfem_ctx* ctx, double bx, double by) {
  return guarded([&] {
    need(ctx != nullptr, "cstfem_add_body_force: NULL context");
    Context& c = ctx->c;
    use_device(&c);
    need(c.nn > 0 && c.have_pattern && c.mesh_pattern, "cstfem_add_body_force: build the mesh pattern first");
    ensure_vectors(c);
    add_body_force_device(c, bx, by);
    return CSTFEM_OK;
  });
}

// bc.cpp:190-253: consistent nodal forces of a traction on one line group
int cstfem_add_traction(cstfem_ctx* ctx, int nedge, const int* edges, int mode, double p0,
                        double p1) {
  return guarded([&] {
    need(ctx && (nedge == 0 || edges), "cstfem_add_traction: NULL argument");
    need(nedge >= 0, "cstfem_add_traction: negative edge count");
    Context& c = ctx->c;
    use_device(&c);
    need(c.nn > 0 && c.have_pattern && c.mesh_pattern, "cstfem_add_traction: build the mesh pattern first");
    need(mode == 0 || mode == 1, "cstfem_add_traction: mode must be 0 (constant) or 1 (kirsch)");
    ensure_vectors(c);
    TractionSource src;
    src.kind = mode == 0 ? TractionSource::kConstant : TractionSource::kKirsch;
    src.p0 = p0;
    src.p1 = p1;
    add_traction_device(c, nedge, edges, src);
    return CSTFEM_OK;
  });
}

int cstfem_add_traction_fn(cstfem_ctx* ctx, int nedge, const int* edges, cstfem_traction_fn fn,
                           void* user) {
  return guarded([&] {
    need(ctx && fn && (nedge == 0 || edges), "cstfem_add_traction_fn: NULL argument");
    need(nedge >= 0, "cstfem_add_traction_fn: negative edge count");
    Context& c = ctx->c;
    use_device(&c);
    need(c.nn > 0 && c.have_pattern && c.mesh_pattern, "cstfem_add_traction_fn: build the mesh pattern first");
    ensure_vectors(c);
    TractionSource src;
    src.kind = TractionSource::kFunction;
    src.fn = fn;
    src.user = user;
    add_traction_device(c, nedge, edges, src);
    return CSTFEM_OK;
  });
}

int cstfem_apply_dirichlet(cstfem_ctx* ctx, int m, const int* dofs, const double* values) {
  return guarded([&] {
    need(ctx && (m == 0 || (dofs && values)), "cstfem_apply_dirichlet: NULL argument");
    need(m >= 0, "cstfem_apply_dirichlet: negative count");
    use_device(&ctx->c);
    apply_dirichlet_device(ctx->c, m, dofs, values);
    return CSTFEM_OK;
  });
}

int cstfem_apply_dirichlet_noalloc(cstfem_ctx* ctx, int m, const int* dofs, const double* values) {
  return guarded([&] {
    need(ctx && (m == 0 || (dofs && values)), "cstfem_apply_dirichlet_noalloc: NULL argument");
    need(m >= 0, "cstfem_apply_dirichlet_noalloc: negative count");
    use_device(&ctx->c);
    apply_dirichlet_noalloc(ctx->c, m, dofs, values);
    return CSTFEM_OK;
  });
}

int cstfem_setup(cstfem_ctx* ctx) {
  return guarded([&] {
    need(ctx != nullptr, "cstfem_setup: NULL context");
    use_device(&ctx->c);
    setup_solver(ctx->c);
    return CSTFEM_OK;
  });
}

int cstfem_solve(cstfem_ctx* ctx, const cstfem_solver_options* opt, cstfem_solver_stats* st,
                 double* x_out, double* history) {
  return guarded([&] {
    need(ctx && opt && st, "cstfem_solve: NULL argument");
    Context& c = ctx->c;
    use_device(&c);
    need(opt->max_iter >= 0, "cstfem_solve: max_iter must be >= 0");
    need(opt->rtol > 0 && std::isfinite(opt->rtol), "cstfem_solve: rtol must be positive");
    std::memset(st, 0, sizeof *st);
    if (!c.setup_done) throw Error(CSTFEM_ERR_ORDER, "cg_jacobi: solve() called before setup()");
    int rc;
    switch (opt->method) {
      case CSTFEM_METHOD_CG_FP64:
        rc = solve_cg_fp64(c, *opt, *st, history);
        break;
      case CSTFEM_METHOD_IR_FP32:
        rc = solve_refined(c, *opt, *st, true);
        break;
      case CSTFEM_METHOD_CG_FP64_REFINED:
        rc = solve_refined(c, *opt, *st, false);
        break;
      case CSTFEM_METHOD_FCG_LEARNED:
        rc = solve_fcg_learned(c, *opt, *st, false);
        break;
      case CSTFEM_METHOD_FCG_LEARNED_FP32:
        rc = solve_fcg_learned(c, *opt, *st, true);
        break;
      default:
        throw Error(CSTFEM_ERR_ARGUMENT, "cstfem_solve: unknown method " + std::to_string(opt->method));
    }
    if (x_out) {
      HIP_CHECK(hipMemcpyAsync(x_out, c.x.get(), sizeof(double) * c.n, hipMemcpyDeviceToHost, c.stream));
      HIP_CHECK(hipStreamSynchronize(c.stream));
    }
    return rc;
  });
}

int cstfem_download_solution(cstfem_ctx* ctx, double* x, int n) {
  return guarded([&] {
    need(ctx && x, "cstfem_download_solution: NULL argument");
    Context& c = ctx->c;
    use_device(&c);
    need(c.x.n > 0 && n == c.n, "cstfem_download_solution: size mismatch or no solution");
    HIP_CHECK(hipMemcpyAsync(x, c.x.get(), sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
    return CSTFEM_OK;
  });
}

int cstfem_spmv(cstfem_ctx* ctx, const double* x, double* y, int n) {
  return guarded([&] {
    need(ctx && x && y, "cstfem_spmv: NULL argument");
    Context& c = ctx->c;
    use_device(&c);
    need(c.have_pattern && n == c.n, "cstfem_spmv: size mismatch or no matrix");
    ensure_vectors(c);
    // staged in the context's work vectors (no allocation per call); the padding row of an
    // odd-sized generic matrix reads zero
    double* dx = c.work.get();
    double* dy = c.Ap.get();
    if (n != 2 * c.nb) HIP_CHECK(hipMemsetAsync(dx, 0, sizeof(double) * 2 * c.nb, c.stream));
    HIP_CHECK(hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    spmv_fp64(c, dx, dy);
    HIP_CHECK(hipMemcpyAsync(y, dy, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
    return CSTFEM_OK;
  });
}

int cstfem_dot(cstfem_ctx* ctx, const double* a, const double* b, int n, double* out) {
  return guarded([&] {
    need(ctx && a && b && out, "cstfem_dot: NULL argument");
    Context& c = ctx->c;
    use_device(&c);
    need(c.have_pattern && n == c.n, "cstfem_dot: size mismatch or no matrix");
    ensure_vectors(c);
    double* da = c.work.get();
    double* db = c.Ap.get();
    if (n != 2 * c.nb) {
      HIP_CHECK(hipMemsetAsync(da, 0, sizeof(double) * 2 * c.nb, c.stream));
      HIP_CHECK(hipMemsetAsync(db, 0, sizeof(double) * 2 * c.nb, c.stream));
    }
    HIP_CHECK(hipMemcpyAsync(da, a, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    HIP_CHECK(hipMemcpyAsync(db, b, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    *out = dot_fp64(c, da, db);
    return CSTFEM_OK;
  });
}

int cstfem_set_material_fields(cstfem_ctx* ctx, const double* E_elem, const double* nu_elem,
                               int plane, double thickness) {
  return guarded([&] {
    need(ctx && E_elem && nu_elem, "cstfem_set_material_fields: NULL argument");
    Context& c = ctx->c;
    use_device(&c);
    need(c.nn > 0, "cstfem_set_material_fields: upload a mesh first");
    need(thickness > 0, "thickness must be positive");
    need(plane == CSTFEM_PLANE_STRESS || plane == CSTFEM_PLANE_STRAIN,
         "plane must be plane_stress or plane_strain");
    for (int64_t e = 0; e < c.ne; ++e) {
      need(E_elem[e] > 0, "element " + std::to_string(e) + ": E must be positive");
      need(nu_elem[e] > -1.0 && nu_elem[e] < 0.5,
           "element " + std::to_string(e) + ": nu must be in (-1, 0.5)");
    }
    c.E_elem.alloc(c.ne);
    c.nu_elem.alloc(c.ne);
    HIP_CHECK(hipMemcpyAsync(c.E_elem.get(), E_elem, sizeof(double) * c.ne, hipMemcpyHostToDevice, c.stream));
    HIP_CHECK(hipMemcpyAsync(c.nu_elem.get(), nu_elem, sizeof(double) * c.ne, hipMemcpyHostToDevice, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
    c.has_E_elem = c.has_nu_elem = true;
    c.thick = thickness;
    c.plane_strain = plane;
    return CSTFEM_OK;
  });
}

int cstfem_batch_define(cstfem_ctx* ctx, int count, const int* node_off) {
  return guarded([&] {
    need(ctx && node_off, "cstfem_batch_define: NULL argument");
    use_device(&ctx->c);
    batch_define(ctx->c, count, node_off);
    return CSTFEM_OK;
  });
}

int cstfem_batch_solve(cstfem_ctx* ctx, const cstfem_solver_options* opt,
                       const cstfem_batch_config* cfg, cstfem_batch_result* results,
                       double* x_out, double* seconds) {
  return guarded([&] {
    need(ctx && opt, "cstfem_batch_solve: NULL argument");
    Context& c = ctx->c;
    use_device(&c);
    need(opt->rtol > 0 && opt->max_iter >= 0, "cstfem_batch_solve: bad rtol or max_iter");
    const cstfem_batch_config def{8, 1, 16};
    std::vector<cstfem_batch_result> res(c.batch_count > 0 ? c.batch_count : 1);
    const double t = batch_solve(c, *opt, cfg ? *cfg : def, res.data());
    if (results) std::copy(res.begin(), res.begin() + c.batch_count, results);
    if (seconds) *seconds = t;
    if (x_out) {
      HIP_CHECK(hipMemcpyAsync(x_out, c.x.get(), sizeof(double) * c.n, hipMemcpyDeviceToHost, c.stream));
      HIP_CHECK(hipStreamSynchronize(c.stream));
    }
    for (int k = 0; k < c.batch_count; ++k)
      if (!res[k].converged) return CSTFEM_NOT_CONVERGED;
    return CSTFEM_OK;
  });
}

int cstfem_batch_harvest(cstfem_ctx* ctx, const cstfem_solver_options* opt, int k_first,
                         int k_last, int* npairs) {
  return guarded([&] {
    need(ctx && opt, "cstfem_batch_harvest: NULL argument");
    need(opt->rtol > 0 && opt->max_iter > 0, "cstfem_batch_harvest: bad rtol or max_iter");
    use_device(&ctx->c);
    const int got = batch_harvest(ctx->c, *opt, k_first, k_last);
    if (npairs) *npairs = got;
    return CSTFEM_OK;
  });
}

int cstfem_batch_autotune(cstfem_ctx* ctx, const cstfem_solver_options* opt,
                          const char* cache_path, cstfem_batch_config* best,
                          double* best_seconds) {
  return guarded([&] {
    need(ctx && opt && best, "cstfem_batch_autotune: NULL argument");
    use_device(&ctx->c);
    *best = batch_autotune(ctx->c, *opt, cache_path, best_seconds);
    return CSTFEM_OK;
  });
}

int cstfem_element_stresses(cstfem_ctx* ctx, double* sig, double* von_mises, double* szz,
                            double* cxy, double* area) {
  return guarded([&] {
    need(ctx && sig, "cstfem_element_stresses: NULL argument");
    use_device(&ctx->c);
    element_stresses(ctx->c, sig, von_mises, szz, cxy, area);
    return CSTFEM_OK;
  });
}

int cstfem_kirsch_metrics(cstfem_ctx* ctx, double sigma_inf, double a, cstfem_kirsch_result* out) {
  return guarded([&] {
    need(ctx && out, "cstfem_kirsch_metrics: NULL argument");
    use_device(&ctx->c);
    kirsch_metrics(ctx->c, sigma_inf, a, out);
    return CSTFEM_OK;
  });
}

int cstfem_mpnn_num_params(void) { return mpnn_num_params(); }

int cstfem_mpnn_init(cstfem_ctx* ctx, int seed) {
  return guarded([&] {
    need(ctx != nullptr, "cstfem_mpnn_init: NULL context");
    use_device(&ctx->c);
    mpnn_init(ctx->c, seed);
    return CSTFEM_OK;
  });
}

int cstfem_mpnn_set_params(cstfem_ctx* ctx, const float* theta, int n) {
  return guarded([&] {
    need(ctx && theta, "cstfem_mpnn_set_params: NULL argument");
    use_device(&ctx->c);
    mpnn_set_params(ctx->c, theta, n);
    return CSTFEM_OK;
  });
}

int cstfem_mpnn_get_params(cstfem_ctx* ctx, float* theta, int n) {
  return guarded([&] {
    need(ctx && theta, "cstfem_mpnn_get_params: NULL argument");
    use_device(&ctx->c);
    mpnn_get_params(ctx->c, theta, n);
    return CSTFEM_OK;
  });
}

int cstfem_mpnn_build(cstfem_ctx* ctx) {
  return guarded([&] {
    need(ctx != nullptr, "cstfem_mpnn_build: NULL context");
    use_device(&ctx->c);
    mpnn_forward(ctx->c);
    return CSTFEM_OK;
  });
}

int cstfem_mpnn_harvest(cstfem_ctx* ctx, int k_first, int k_last, int stride, int* npairs) {
  return guarded([&] {
    need(ctx != nullptr, "cstfem_mpnn_harvest: NULL context");
    use_device(&ctx->c);
    const int got = mpnn_harvest(ctx->c, k_first, k_last, stride);
    if (npairs) *npairs = got;
    return CSTFEM_OK;
  });
}

int cstfem_mpnn_add_white_pairs(cstfem_ctx* ctx, int count, int seed) {
  return guarded([&] {
    need(ctx != nullptr, "cstfem_mpnn_add_white_pairs: NULL context");
    use_device(&ctx->c);
    mpnn_add_white_pairs(ctx->c, count, seed);
    return CSTFEM_OK;
  });
}

int cstfem_mpnn_project_pairs(cstfem_ctx* ctx) {
  return guarded([&] {
    need(ctx != nullptr, "cstfem_mpnn_project_pairs: NULL context");
    use_device(&ctx->c);
    mpnn_project_pairs(ctx->c);
    return CSTFEM_OK;
  });
}

int cstfem_mpnn_set_pairs(cstfem_ctx* ctx, int K, const float* R, const float* E) {
  return guarded([&] {
    need(ctx != nullptr, "cstfem_mpnn_set_pairs: NULL context");
    need(K > 0, "cstfem_mpnn_set_pairs: K must be positive");
    use_device(&ctx->c);
    mpnn_set_pairs(ctx->c, K, R, E);
    return CSTFEM_OK;
  });
}

int cstfem_mpnn_loss_grad(cstfem_ctx* ctx, double* loss) {
  return guarded([&] {
    need(ctx && loss, "cstfem_mpnn_loss_grad: NULL argument");
    use_device(&ctx->c);
    *loss = mpnn_loss_grad(ctx->c);
    return CSTFEM_OK;
  });
}

int cstfem_mpnn_train_step(cstfem_ctx* ctx, float lr, double* loss) {
  return guarded([&] {
    need(ctx != nullptr, "cstfem_mpnn_train_step: NULL context");
    need(lr > 0.f && std::isfinite(lr), "cstfem_mpnn_train_step: lr must be positive");
    use_device(&ctx->c);
    const double l = mpnn_train_step(ctx->c, lr);
    if (loss) *loss = l;
    return CSTFEM_OK;
  });
}

int cstfem_mpnn_download(cstfem_ctx* ctx, int what, float* out, int64_t count) {
  return guarded([&] {
    need(ctx && out, "cstfem_mpnn_download: NULL argument");
    use_device(&ctx->c);
    mpnn_download(ctx->c, what, out, count);
    return CSTFEM_OK;
  });
}

int cstfem_mg_setup(cstfem_ctx* ctx, int max_levels, int coarse_max, int fine_smoother,
                    int kcycle_levels) {
  return guarded([&] {
    need(ctx != nullptr, "cstfem_mg_setup: NULL context");
    use_device(&ctx->c);
    mg_setup(ctx->c, max_levels, coarse_max, fine_smoother, kcycle_levels);
    return CSTFEM_OK;
  });
}

int cstfem_mg_prepare(cstfem_ctx* ctx) {
  return guarded([&] {
    need(ctx != nullptr, "cstfem_mg_prepare: NULL context");
    use_device(&ctx->c);
    mg_prepare(ctx->c);
    return CSTFEM_OK;
  });
}

int cstfem_mg_levels(cstfem_ctx* ctx, int* levels) {
  return guarded([&] {
    need(ctx && levels, "cstfem_mg_levels: NULL argument");
    *levels = mg_levels(ctx->c);
    return CSTFEM_OK;
  });
}

int cstfem_mg_level_info(cstfem_ctx* ctx, int level, int* n, int* nagg, int64_t* nnz,
                         double* lam) {
  return guarded([&] {
    need(ctx != nullptr, "cstfem_mg_level_info: NULL context");
    const int64_t z = mg_level_info(ctx->c, level, n, nagg, lam);
    need(z >= 0, "cstfem_mg_level_info: no level " + std::to_string(level));
    if (nnz) *nnz = z;
    return CSTFEM_OK;
  });
}

int cstfem_mg_tail(cstfem_ctx* ctx, int* first_level, int* packed) {
  return guarded([&] {
    need(ctx != nullptr, "cstfem_mg_tail: NULL context");
    mg_tail_info(ctx->c, first_level, packed);
    return CSTFEM_OK;
  });
}

int cstfem_mg_download_aggregates(cstfem_ctx* ctx, int level, int* agg, int count) {
  return guarded([&] {
    need(ctx && agg, "cstfem_mg_download_aggregates: NULL argument");
    mg_download_agg(ctx->c, level, agg, count);
    return CSTFEM_OK;
  });
}

int cstfem_mg_download_matrix(cstfem_ctx* ctx, int level, int* row_ptr, int* cols, double* vals) {
  return guarded([&] {
    need(ctx && row_ptr && cols && vals, "cstfem_mg_download_matrix: NULL argument");
    use_device(&ctx->c);
    mg_download_matrix(ctx->c, level, row_ptr, cols, vals);
    return CSTFEM_OK;
  });
}

int cstfem_mg_apply(cstfem_ctx* ctx, const double* r, double* z, int n) {
  return guarded([&] {
    need(ctx && r && z, "cstfem_mg_apply: NULL argument");
    Context& c = ctx->c;
    use_device(&c);
    need(n == c.n, "cstfem_mg_apply: size mismatch");
    DBuf<double> dr, dz;
    dr.alloc(n);
    dz.alloc(n);
    HIP_CHECK(hipMemcpyAsync(dr.get(), r, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    mg_enqueue_apply(c, dr.get(), dz.get(), nullptr);
    HIP_CHECK(hipMemcpyAsync(z, dz.get(), sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
    return CSTFEM_OK;
  });
}

int cstfem_profile_kernel(cstfem_ctx* ctx, int kind, int reps, int64_t bytes, double* avg_seconds) {
  return guarded([&] {
    need(ctx && avg_seconds, "cstfem_profile_kernel: NULL argument");
    need(kind >= 0 && kind <= 10, "cstfem_profile_kernel: kind must be 0..10");
    need(reps > 0, "cstfem_profile_kernel: reps must be positive");
    need((kind != 4 && kind != 10) || bytes >= 32,
         "cstfem_profile_kernel: copy and read need bytes >= 32");
    use_device(&ctx->c);
    *avg_seconds = profile_kernel(ctx->c, kind, reps, bytes);
    return CSTFEM_OK;
  });
}

}  // extern "C"
