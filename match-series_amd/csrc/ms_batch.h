// Several independent series in one call (msg_match_series_batch). No counterpart in the reference:
// its program matches one series per run (matchSeries.h:1249-1320); each series here follows
// SeriesMatching::matchAndAverageSeries (matchSeries.h:87-269) and averageSeries (L353-451) exactly
// as msg_match_series does, and its results are bitwise those of msg_match_series on it alone.
//
// Scheduling. Series are grouped by frame size and solve parameters (one group = one context shape
// and one set of level-solve parameters). Within a group:
//   * the pairwise solves of all its series are batched together, across series, in contexts of up
//     to `batch` pairs (every pair's arithmetic is independent of the batch it sits in), in the order
//     of their template index, on a second host thread;
//   * the accumulate-and-refine chains advance in lockstep, refinement i of every series that has a
//     template i in one batched solve (a group of k series pays one chain's launches for k chains),
//     each step as soon as the pairwise solves up to template i are done;
//   * the series' reconstructions follow.
// Groups run concurrently, each on its own host thread and streams, so groups of different frame
// sizes share the GPU (a latency-bound chain of one size beside the batches of another).
// Included by ms_gpu.hip inside its anonymous namespace, after the series implementation.

// A persistent host thread per concurrent group (thread_local scratch buffers stay with their
// thread, so a worker is reused rather than recreated for every call).
constexpr int BATCH_WORKERS = 8;
Worker *batch_worker(int k) {  // k < 2 * BATCH_WORKERS: a group's chain thread, then its pairwise thread
  static Worker *w[2 * BATCH_WORKERS] = {};
  static std::mutex mu;
  std::lock_guard<std::mutex> g(mu);
  if (!w[k]) w[k] = new Worker;
  return w[k];
}

// Parameters that must agree for two series to share contexts and level solves.
bool same_group(const msg_series_params &a, const msg_series_params &b) {
  return a.start_level == b.start_level && a.stop_level == b.stop_level &&
         a.refine_start_level == b.refine_start_level && a.refine_stop_level == b.refine_stop_level &&
         a.precision_level == b.precision_level && a.lambda == b.lambda && a.lambda_factor == b.lambda_factor &&
         a.max_iter == b.max_iter && a.stop_eps == b.stop_eps && a.saturation_pct == b.saturation_pct &&
         a.normalize_min_to_zero == b.normalize_min_to_zero && a.dont_normalize == b.dont_normalize &&
         a.resample_prolong == b.resample_prolong && a.dont_accumulate == b.dont_accumulate && a.spec == b.spec &&
         a.pre_smooth_sigma == b.pre_smooth_sigma && a.pairwise_batch == b.pairwise_batch &&
         a.data_term == b.data_term && a.y_reg_factor == b.y_reg_factor &&
         a.average_ext_mean == b.average_ext_mean && a.average_nearest == b.average_nearest;
}

// What msg_match_series_batch does not take (each has its own entry point or context call).
int check_batch_params(const msg_series_params *p) {
  if (p->reduce_deformations) return fail(MSG_ERR_ARG, "batch: reduce_deformations is msg_match_series only");
  if (p->corr_init) return fail(MSG_ERR_ARG, "batch: useCorrelationToInitTranslation is msg_match_series only");
  if (p->calc_inverse) return fail(MSG_ERR_ARG, "batch: calcInverseDeformation is msg_match_series only");
  if (p->alt_start_level > 0) return fail(MSG_ERR_ARG, "batch: altStartLevel is msg_match_series only");
  if (p->pair0_init || p->raw_templates) return fail(MSG_ERR_ARG, "batch: pair0_init / raw_templates unsupported");
  if (p->keep_start_transformation) return fail(MSG_ERR_ARG, "batch: keep_start_transformation unsupported");
  if (p->data_term == MSG_XSHIFT) return fail(MSG_ERR_ARG, "batch: MSG_XSHIFT series are msg_match_series only");
  if (p->precision_level < 1 || p->precision_level > 13) return fail(MSG_ERR_ARG, "batch: precision_level 1..13");
  if (p->start_level < 1 || p->stop_level > p->precision_level || p->start_level > p->stop_level)
    return fail(MSG_ERR_ARG, "batch: invalid level range");
  if (!p->dont_accumulate && (p->refine_start_level < 1 || p->refine_stop_level > p->precision_level ||
                              p->refine_start_level > p->refine_stop_level))
    return fail(MSG_ERR_ARG, "batch: invalid refine level range");
  return MSG_OK;
}

// Pairwise solves per batch: all of a group's pairs at small sizes, fewer at large ones (a 2048^2
// FP32 pair's context holds about 250 MB).
int batch_pairs_per_context(const msg_series_params *p, int total) {
  if (p->pairwise_batch > 0) return std::max(1, std::min(p->pairwise_batch, total));
  const int L = p->precision_level;
  const int cap = L <= 8 ? 64 : (L == 9 ? 32 : (L == 10 ? 16 : 8));
  return std::max(1, std::min(cap, total));
}

template <typename S>
int batch_group(const std::vector<const msg_series_job *> &jobs, int slot) {
  const msg_series_params *p = jobs[0]->params;
  const int L = p->precision_level, N = 1 << L;
  const size_t nn = (size_t)N * N;
  const int dt = sizeof(S) == 8 ? MSG_F64 : MSG_F32;
  const int J = (int)jobs.size();
  msg_solve_params sp;
  sp.lambda = p->lambda;
  sp.lambda_factor = p->lambda_factor;
  sp.lambda_ref_level = p->stop_level;
  sp.max_iter = p->max_iter;
  sp.stop_eps = p->stop_eps;
  sp.resample_prolong = p->resample_prolong;
  sp.spec = p->spec;
  struct StreamGuard {
    hipStream_t s = nullptr;
    ~StreamGuard() {
      if (s) (void)hipStreamDestroy(s);
    }
  } sg;
  HIPCHK(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
  hipStream_t s = sg.s;
  auto tpl = [&](int j, int i) { return (const S *)jobs[j]->templates + (size_t)i * nn; };
  auto prepare = [&](msg_ctx *c, int nb) -> int {  // normalise / pre-smooth the level-L images, pyramids
    if (!p->dont_normalize) {
      CHK(normalize_impl<S>(c->R[L], N, nb, p->saturation_pct, p->normalize_min_to_zero, c->stream));
      CHK(normalize_impl<S>(c->T[L], N, nb, p->saturation_pct, p->normalize_min_to_zero, c->stream));
    }
    if (p->pre_smooth_sigma > 0) {
      const double sig = p->pre_smooth_sigma / (double)(N - 1);
      CHK(h1_impl<S>(c->R[L], c->R[L], N, nb, sig, c->stream));
      CHK(h1_impl<S>(c->T[L], c->T[L], N, nb, sig, c->stream));
    }
    CHK(rebuild_pyramid<S>(c, c->R, -1, 1));
    CHK(rebuild_pyramid<S>(c, c->T, -1, 1));
    return MSG_OK;
  };

  // --- pairwise solves of every series, batched across series, in template order ---
  std::vector<DevBuf> pw(J);
  std::vector<std::pair<int, int>> pairs;  // (series, template), ordered by template
  int maxn = 0;
  for (int j = 0; j < J; ++j) {
    HIPCHK(hipMalloc(&pw[j].p, sizeof(S) * 2 * nn * jobs[j]->n));
    for (int i = 0; i < jobs[j]->n; ++i) jobs[j]->energies[i] = NAN;
    maxn = std::max(maxn, jobs[j]->n);
  }
  for (int i = 0; i < maxn; ++i)
    for (int j = 0; j < J; ++j)
      if (jobs[j]->n > i) pairs.push_back({j, i});
  const int total = (int)pairs.size();
  const int b = batch_pairs_per_context(p, total);
  auto pairwise = [&](int k0, int nb) -> int {
    msg_ctx *pc = nullptr;
    CHK(ctx_acquire(L, nb, dt, &pc));
    CtxGuard gpc{pc};
    CHK(msg_ctx_set_data_term(pc, p->data_term));
    hipStream_t cs = pc->stream;
    for (int k = 0; k < nb; ++k) {
      const int j = pairs[k0 + k].first, i = pairs[k0 + k].second;
      const S *r = (i == 0 || p->dont_accumulate) ? (const S *)jobs[j]->reference : tpl(j, i - 1);
      HIPCHK(hipMemcpyAsync((S *)pc->R[L] + k * nn, r, nn * sizeof(S), hipMemcpyDeviceToDevice, cs));
      HIPCHK(hipMemcpyAsync((S *)pc->T[L] + k * nn, tpl(j, i), nn * sizeof(S), hipMemcpyDeviceToDevice, cs));
    }
    CHK(prepare(pc, nb));
    for (int l = 0; l <= L; ++l)
      HIPCHK(hipMemsetAsync(pc->phi[l], 0, sizeof(S) * ((size_t)2 << (2 * l)) * nb, cs));
    std::vector<double> e(nb);
    CHK(solve_impl<S>(pc, p->start_level, p->stop_level, &sp, e.data()));
    for (int k = 0; k < nb; ++k) {
      const int j = pairs[k0 + k].first, i = pairs[k0 + k].second;
      if (i == 0 || p->dont_accumulate) jobs[j]->energies[i] = e[k];  // refinements replace the others
      HIPCHK(hipMemcpyAsync((S *)pw[j].p + (size_t)i * 2 * nn, (const S *)pc->phi[L] + (size_t)k * 2 * nn,
                            sizeof(S) * 2 * nn, hipMemcpyDeviceToDevice, cs));
    }
    HIPCHK(hipStreamSynchronize(cs));
    return MSG_OK;
  };
  Progress pr;  // pr.done: pairs finished, a prefix of `pairs`
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  batch_worker(BATCH_WORKERS + slot)->post([&, dev] {
    int rc = hipSetDevice(dev) == hipSuccess ? MSG_OK : fail(MSG_ERR_HIP, "hipSetDevice failed");
    for (int k0 = 0; k0 < total && rc == MSG_OK; k0 += b) {
      {
        std::lock_guard<std::mutex> g(pr.mu);
        if (pr.abort) break;
      }
      const int nb = std::min(b, total - k0);
      rc = pairwise(k0, nb);
      std::lock_guard<std::mutex> g(pr.mu);
      if (rc == MSG_OK) pr.done = k0 + nb;
      pr.cv.notify_all();
    }
    std::lock_guard<std::mutex> g(pr.mu);
    if (rc != MSG_OK) {
      pr.rc = rc;
      pr.err = g_err;
    }
    pr.finished = true;
    pr.cv.notify_all();
  });
  // Waits until the first `count` pairs are solved; false after a failure.
  auto wait_pairs = [&](int count) {
    std::unique_lock<std::mutex> l(pr.mu);
    pr.cv.wait(l, [&] { return pr.done >= count || pr.rc != MSG_OK || pr.finished; });
    return pr.done >= count;
  };
  // Stops the pairwise thread (after an error) and waits for it to let go of the buffers.
  auto join = [&](int code) {
    std::unique_lock<std::mutex> l(pr.mu);
    if (code != MSG_OK) pr.abort = true;
    pr.cv.wait(l, [&] { return pr.finished; });
    return code;
  };

  // --- the chains in lockstep: refinement i of every series with a template i in one solve ---
  const int chain_rc = [&]() -> int {
    if (p->dont_accumulate) {
      if (!wait_pairs(total)) return pr.rc != MSG_OK ? pr.rc : MSG_ERR_STATE;
      for (int j = 0; j < J; ++j)
        HIPCHK(hipMemcpyAsync(jobs[j]->deformations, pw[j].p, sizeof(S) * 2 * nn * jobs[j]->n,
                              hipMemcpyDeviceToDevice, s));
      return MSG_OK;
    }
    std::vector<DevBuf> acc(J), refn(J);
    DevBuf prev;
    HIPCHK(hipMalloc(&prev.p, sizeof(S) * 2 * nn));
    for (int j = 0; j < J; ++j) {
      HIPCHK(hipMalloc(&acc[j].p, sizeof(S) * 2 * nn));
      HIPCHK(hipMemsetAsync(acc[j].p, 0, sizeof(S) * 2 * nn, s));
      // the reference as the chain's context prepares it
      HIPCHK(hipMalloc(&refn[j].p, sizeof(S) * nn));
      HIPCHK(hipMemcpyAsync(refn[j].p, jobs[j]->reference, nn * sizeof(S), hipMemcpyDeviceToDevice, s));
      if (!p->dont_normalize)
        CHK(normalize_impl<S>(refn[j].p, N, 1, p->saturation_pct, p->normalize_min_to_zero, s));
      if (p->pre_smooth_sigma > 0)
        CHK(h1_impl<S>(refn[j].p, refn[j].p, N, 1, p->pre_smooth_sigma / (double)(N - 1), s));
    }
    int upto = 0;  // pairs with template index <= i (a prefix of `pairs`)
    for (int i = 0; i < maxn; ++i) {
      std::vector<int> act;
      for (int j = 0; j < J; ++j)
        if (jobs[j]->n > i) act.push_back(j);
      upto += (int)act.size();
      if (!wait_pairs(upto)) return pr.rc != MSG_OK ? pr.rc : MSG_ERR_STATE;
      for (int j : act) {
        HIPCHK(hipMemcpyAsync(prev.p, acc[j].p, sizeof(S) * 2 * nn, hipMemcpyDeviceToDevice, s));
        CHK(compose_impl<S>((const S *)pw[j].p + (size_t)i * 2 * nn, prev.p, acc[j].p, N, s));
      }
      HIPCHK(hipStreamSynchronize(s));
      if (i > 0) {
        const int nb = (int)act.size();
        msg_ctx *rc = nullptr;
        CHK(ctx_acquire(L, nb, dt, &rc));
        CtxGuard grc{rc};
        CHK(msg_ctx_set_data_term(rc, p->data_term));
        hipStream_t cs = rc->stream;
        for (int k = 0; k < nb; ++k) {
          const int j = act[k];
          HIPCHK(hipMemcpyAsync((S *)rc->R[L] + k * nn, refn[j].p, nn * sizeof(S), hipMemcpyDeviceToDevice, cs));
          HIPCHK(hipMemcpyAsync((S *)rc->T[L] + k * nn, tpl(j, i), nn * sizeof(S), hipMemcpyDeviceToDevice, cs));
          HIPCHK(hipMemcpyAsync((S *)rc->phi[L] + (size_t)k * 2 * nn, acc[j].p, sizeof(S) * 2 * nn,
                                hipMemcpyDeviceToDevice, cs));
        }
        // the reference is already prepared; only the templates are normalised here
        if (!p->dont_normalize)
          CHK(normalize_impl<S>(rc->T[L], N, nb, p->saturation_pct, p->normalize_min_to_zero, cs));
        if (p->pre_smooth_sigma > 0)
          CHK(h1_impl<S>(rc->T[L], rc->T[L], N, nb, p->pre_smooth_sigma / (double)(N - 1), cs));
        CHK(rebuild_pyramid<S>(rc, rc->R, -1, 1));
        CHK(rebuild_pyramid<S>(rc, rc->T, -1, 1));
        CHK(rebuild_pyramid<S>(rc, rc->phi, -1, 2));
        std::vector<double> e(nb);
        CHK(solve_impl<S>(rc, p->refine_start_level, p->refine_stop_level, &sp, e.data()));
        for (int k = 0; k < nb; ++k) {
          const int j = act[k];
          jobs[j]->energies[i] = e[k];
          HIPCHK(hipMemcpyAsync(acc[j].p, (const S *)rc->phi[L] + (size_t)k * 2 * nn, sizeof(S) * 2 * nn,
                                hipMemcpyDeviceToDevice, cs));
        }
        HIPCHK(hipStreamSynchronize(cs));
      }
      for (int j : act)
        HIPCHK(hipMemcpyAsync((S *)jobs[j]->deformations + (size_t)i * 2 * nn, i > 0 ? acc[j].p : pw[j].p,
                              sizeof(S) * 2 * nn, hipMemcpyDeviceToDevice, s));
    }
    return MSG_OK;
  }();
  join(chain_rc);
  if (chain_rc != MSG_OK) return chain_rc == pr.rc && !pr.err.empty() ? fail(chain_rc, pr.err) : chain_rc;
  if (pr.rc != MSG_OK) return fail(pr.rc, pr.err);

  // --- reconstructions (averageSeries, matchSeries.h:353-451) ---
  for (int j = 0; j < J; ++j) {
    const int n = jobs[j]->n;
    const S *t = (const S *)jobs[j]->templates;
    const S *defs = (const S *)jobs[j]->deformations;
    DevBuf stack;
    HIPCHK(hipMalloc(&stack.p, sizeof(S) * nn * n));
    const int nearest = p->average_nearest ? 1 : 0;
    if (p->average_ext_mean) {
      for (int i = 0; i < n; ++i) {
        double ext = 0;
        CHK(mean_value<S>(t + (size_t)i * nn, (int)nn, &ext, s));
        hipLaunchKernelGGL((msg::k_warp<S>), dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, t + (size_t)i * nn,
                           defs + (size_t)i * 2 * nn, (S *)stack.p + (size_t)i * nn, N, 1, 0, 0, ext, nearest);
        HIPCHK(hipGetLastError());
      }
    } else {
      const size_t cnt = nn * n;
      hipLaunchKernelGGL((msg::k_warp<S>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, t, defs,
                         (S *)stack.p, N, n, 0, 0, (double)INFINITY, nearest);
      HIPCHK(hipGetLastError());
    }
    double fill = 0;
    CHK(mean_value<S>(t + (size_t)(n - 1) * nn, (int)nn, &fill, s));
    CHK(stats_impl<S>(stack.p, n, N, fill, jobs[j]->mean, jobs[j]->count, jobs[j]->median, s));
    HIPCHK(hipStreamSynchronize(s));  // the stack is freed at the end of the iteration
  }
  return MSG_OK;
}

template <typename S>
int match_series_batch_impl(int njobs, const msg_series_job *jobs) {
  // groups in order of first appearance
  std::vector<std::vector<const msg_series_job *>> groups;
  for (int k = 0; k < njobs; ++k) {
    bool placed = false;
    for (auto &g : groups)
      if (same_group(*g[0]->params, *jobs[k].params)) {
        g.push_back(&jobs[k]);
        placed = true;
        break;
      }
    if (!placed) groups.push_back({&jobs[k]});
  }
  struct Done {
    std::mutex mu;
    std::condition_variable cv;
    int left = 0, rc = MSG_OK;
    std::string err;
  } done;
  done.left = (int)groups.size();
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  for (size_t g = 0; g < groups.size(); ++g) {
    const std::vector<const msg_series_job *> *grp = &groups[g];
    const int slot = (int)(g % BATCH_WORKERS);
    batch_worker(slot)->post([grp, dev, slot, &done] {
      int rc = hipSetDevice(dev) == hipSuccess ? MSG_OK : fail(MSG_ERR_HIP, "hipSetDevice failed");
      if (rc == MSG_OK) rc = batch_group<S>(*grp, slot);
      std::lock_guard<std::mutex> l(done.mu);
      if (rc != MSG_OK && done.rc == MSG_OK) {
        done.rc = rc;
        done.err = g_err;
      }
      if (--done.left == 0) done.cv.notify_all();
    });
  }
  std::unique_lock<std::mutex> l(done.mu);
  done.cv.wait(l, [&] { return done.left == 0; });
  if (done.rc != MSG_OK) return fail(done.rc, done.err);
  return MSG_OK;
}
