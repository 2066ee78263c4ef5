// Lockstep batch driver: many independent ALS_SPD + round problems advance together, and every step of every
// problem (a stack push, a local right-hand side, a local operator product, a CG step, a QR or RQ move, a pivoted
// QC, a Jacobi SVD) is one grouped launch for all problems (group.hip and the grouped GEMM, Householder and Jacobi
// kernels). Shapes differ from problem to problem (uneven d, n and ranks); each launch carries a task table.
// Host decisions (ranks, convergence, energies, the stop rule) are read back once per step for all problems.
//
// The per-problem algorithm is the one of als_solve (one site, SPD operator, FP64 local CG, or FP32 CG inside FP64
// iterative refinement, with the same tolerances, iteration caps and zero-start rule) and round (QC sweep to the
// last core, then round_edge from the right with the SVD truncation rule). Trajectories agree with the per-problem
// drivers to the local solver's tolerance; rank decisions are made by the same kernels and rules.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <functional>
#include <vector>

#include "tt.hpp"

namespace xtt {

namespace {

const std::vector<size_t> kAm{2, 0, 3, 1};   // operator core (p,n,n',q) → (n,q,p,n'): left pushes, local operator
const std::vector<size_t> kAmM{1, 0, 3, 2};  // → (n,p,q,n'): right pushes (the left push on the mirrored site)

struct Lock {
	const TT* A = nullptr;
	TT* x = nullptr;
	const TT* b = nullptr;
	size_t d = 0;
	std::vector<DTensor> opL, opR, rhsL, rhsR, am, amM;
	size_t first = 0, second = 0, curr = 0;
	bool increasing = true, active = false;
	double le2 = 1e102, le = 1e101, en = 1e100;
	size_t halfSweeps = 0;      // of the running ALS call
	size_t halfSweepsAll = 0;   // over all calls
	bool canonAtEnd = true;
	size_t corePosAtEnd = 0;
	size_t localSolves = 0, cgIters = 0, refinements = 0;
};

DTensor ones_dev(const dims_t& dims, std::vector<MapTask>& maps) {
	DTensor t(dims);
	maps.push_back(map_fill(t.data(), 1.0, t.size()));
	return t;
}

GemmTask gt(const GemmDesc& d, const double* A, const double* B, double* C) { return GemmTask{d, A, B, C}; }

GemmDesc stage2(size_t a, size_t n, size_t c, size_t P, size_t Q, const int* skip) {
	GemmDesc g = gemm_desc(0, 0, n * Q, c, P * n);
	g.batch = a;
	g.sA = 0;
	g.sB = (long long)(P * n * c);
	g.sC = (long long)(n * Q * c);
	g.skip = skip;
	return g;
}

// ---- prepare_x_for_als (als.cpp:330-407): full-rank boundary cores become identities of their unfoldings, their
// content absorbed into the neighbour; one grouped GEMM and one grouped identity fill per absorbed core ------------
struct Absorb {
	size_t j;
	DTensor id, t;
};
void prepare_x_all(std::vector<Lock>& P, hipStream_t s) {
	const size_t J = P.size();
	std::vector<size_t> f(J, 0), l(J), dimF(J, 1), dimL(J, 1);
	std::vector<char> goF(J, 1), goL(J, 1);
	for (size_t j = 0; j < J; ++j) l[j] = P[j].d;
	for (;;) {  // from the left
		std::vector<GemmTask> g;
		std::vector<MapTask> maps;
		std::vector<Absorb> upd;
		for (size_t j = 0; j < J; ++j) {
			TT& x = *P[j].x;
			if (!goF[j] || f[j] + 1 >= P[j].d) { goF[j] = 0; continue; }
			const size_t localDim = x.dims[f[j]], newProd = dimF[j] * localDim;
			if (x.rank(f[j]) < newProd) { goF[j] = 0; continue; }
			const DTensor& cur = x.cores[f[j]];        // (dimProd·n) × newProd
			const DTensor& nx = x.cores[f[j] + 1];     // newProd × (n'·r'')
			DTensor t({cur.dims[0] * cur.dims[1], nx.dims[1], nx.dims[2]});
			g.push_back(gt(gemm_desc(0, 0, cur.dims[0] * cur.dims[1], nx.dims[1] * nx.dims[2], cur.dims[2]), cur.data(), nx.data(), t.data()));
			DTensor id({dimF[j], localDim, newProd});
			maps.push_back(map_eye(id.data(), newProd));
			upd.push_back(Absorb{j, std::move(id), std::move(t)});
			dimF[j] = newProd;
		}
		if (upd.empty()) break;
		gemm_group_f64(g, s);
		map_group(maps, s);
		for (Absorb& u : upd) {
			TT& x = *P[u.j].x;
			x.set_component(f[u.j] + 1, std::move(u.t));
			x.set_component(f[u.j], std::move(u.id));
			f[u.j] += 1;
		}
	}
	for (;;) {  // from the right
		std::vector<GemmTask> g;
		std::vector<MapTask> maps;
		std::vector<Absorb> upd;
		for (size_t j = 0; j < J; ++j) {
			TT& x = *P[j].x;
			if (!goL[j] || l[j] <= f[j] + 1) { goL[j] = 0; continue; }
			const size_t localDim = x.dims[l[j] - 1], newProd = dimL[j] * localDim;
			if (x.rank(l[j] - 2) < newProd) { goL[j] = 0; continue; }
			const DTensor& cur = x.cores[l[j] - 1];    // newProd × (n·dimProd)
			const DTensor& pv = x.cores[l[j] - 2];     // (r''·n') × newProd
			DTensor t({pv.dims[0], pv.dims[1], cur.dims[1] * cur.dims[2]});
			g.push_back(gt(gemm_desc(0, 0, pv.dims[0] * pv.dims[1], cur.dims[1] * cur.dims[2], cur.dims[0]), pv.data(), cur.data(), t.data()));
			DTensor id({newProd, localDim, dimL[j]});
			maps.push_back(map_eye(id.data(), newProd));
			upd.push_back(Absorb{j, std::move(id), std::move(t)});
			dimL[j] = newProd;
		}
		if (upd.empty()) break;
		gemm_group_f64(g, s);
		map_group(maps, s);
		for (Absorb& u : upd) {
			TT& x = *P[u.j].x;
			x.set_component(l[u.j] - 2, std::move(u.t));
			x.set_component(l[u.j] - 1, std::move(u.id));
			l[u.j] -= 1;
		}
	}
	for (size_t j = 0; j < J; ++j) {
		Lock& L = P[j];
		TT& x = *L.x;
		if (L.canonAtEnd && L.corePosAtEnd < f[j]) {
			x.corePosition = f[j]; x.canonicalized = true;
		} else {
			if (L.canonAtEnd && L.corePosAtEnd >= l[j]) { x.corePosition = l[j] - 1; x.canonicalized = true; }
			if (x.corePosition != f[j] || !x.canonicalized) move_core(x, f[j], true);
		}
		L.first = f[j];
		L.second = l[j];
	}
}

// ---- grouped stack pushes -----------------------------------------------------------------------------------
struct Push {
	Lock* L;
	size_t i;
	bool left;
	DTensor Xr, op, rb;
	DBuf<double> T1, T2, Y;
};

void pushes(std::vector<Push>& P, hipStream_t s) {
	if (P.empty()) return;
	ScratchFrame frame;  // Xr, T1, T2, Y
	std::vector<MapTask> maps;
	for (Push& p : P) {
		if (p.left) continue;
		const DTensor& X = p.L->x->cores[p.i];
		p.Xr = DTensor::scratch({X.dims[2], X.dims[1], X.dims[0]});
		const size_t sh[3] = {2, 1, 0};
		maps.push_back(map_permute(3, X.dims.data(), sh, X.data(), p.Xr.data()));
	}
	map_group(maps, s);
	std::vector<GemmTask> g1, g2, g3;
	for (Push& p : P) {
		Lock& L = *p.L;
		const DTensor& X = L.x->cores[p.i];
		const DTensor& Xe = p.left ? X : p.Xr;
		const DTensor& St = p.left ? L.opL.back() : L.opR.back();
		const DTensor& Am = p.left ? L.am[p.i] : L.amM[p.i];
		const DTensor& bc = L.b->cores[p.i];
		const size_t a = Xe.dims[0], n = Xe.dims[1], c = Xe.dims[2], Pp = Am.dims[2], Q = Am.dims[1];
		p.T1.scratch(a * Pp * n * c);
		p.T2.scratch(a * n * Q * c);
		p.op = DTensor({c, Q, c});
		g1.push_back(gt(gemm_desc(0, 0, a * Pp, n * c, a), St.data(), Xe.data(), p.T1.get()));
		g2.push_back(gt(stage2(a, n, c, Pp, Q, nullptr), Am.data(), p.T1.get(), p.T2.get()));
		g3.push_back(gt(gemm_desc(1, 0, c, Q * c, a * n), Xe.data(), p.T2.get(), p.op.data()));
		// right-hand-side stack (on the unmirrored core X = (a0, n, c0))
		const size_t a0 = X.dims[0], c0 = X.dims[2], beta = bc.dims[0], beta2 = bc.dims[2];
		if (p.left) {
			const DTensor& Lb = L.rhsL.back();
			p.Y.scratch(beta * n * c0);
			p.rb = DTensor({beta2, c0});
			g1.push_back(gt(gemm_desc(0, 0, beta, n * c0, a0), Lb.data(), X.data(), p.Y.get()));
			g2.push_back(gt(gemm_desc(1, 0, beta2, c0, beta * n), bc.data(), p.Y.get(), p.rb.data()));
		} else {
			const DTensor& Rb = L.rhsR.back();
			p.Y.scratch(n * beta2 * a0);
			p.rb = DTensor({beta, a0});
			GemmDesc z = gemm_desc(0, 1, beta2, a0, c0);
			z.ldb = n * c0; z.sA = 0; z.sB = (long long)c0; z.sC = (long long)(beta2 * a0); z.batch = n;
			g1.push_back(gt(z, Rb.data(), X.data(), p.Y.get()));
			g2.push_back(gt(gemm_desc(0, 0, beta, a0, n * beta2), bc.data(), p.Y.get(), p.rb.data()));
		}
	}
	gemm_group_f64(g1, s);
	gemm_group_f64(g2, s);
	gemm_group_f64(g3, s);
	for (Push& p : P) {
		Lock& L = *p.L;
		if (p.left) { L.opL.push_back(std::move(p.op)); L.rhsL.push_back(std::move(p.rb)); }
		else { L.opR.push_back(std::move(p.op)); L.rhsR.push_back(std::move(p.rb)); }
	}
}

// ---- the local problem of a site: right-hand side and matrix-free operator -----------------------------------
struct Site {
	Lock* L;
	size_t a, n, c, P, Q;
	const DTensor *Lst, *Rst, *Am;
	DTensor bl;              // local right-hand side (a, n, c)
	DBuf<double> T1, T2, Tb;
};
void site_init(Site& S, Lock* L) {
	S.L = L;
	const size_t k = L->curr;
	const DTensor& X = L->x->cores[k];
	S.a = X.dims[0]; S.n = X.dims[1]; S.c = X.dims[2];
	S.Lst = &L->opL.back(); S.Rst = &L->opR.back(); S.Am = &L->am[k];
	S.P = S.Am->dims[2]; S.Q = S.Am->dims[1];
	S.T1.scratch(S.a * S.P * S.n * S.c);
	S.T2.scratch(S.a * S.n * S.Q * S.c);
	S.bl = DTensor::scratch({S.a, S.n, S.c});
}
// y = A_loc·v for every site, three grouped launches (the first two also compute the local right-hand sides
// when rhs is set); skip[j] (optional): site j's products are skipped while *skip[j] != 0
// (sel: only the listed sites)
void local_apply(std::vector<Site>& S, const std::vector<const double*>& v, const std::vector<double*>& y, bool rhs,
                 const std::vector<const int*>* skip, hipStream_t s, const std::vector<size_t>* sel = nullptr) {
	std::vector<GemmTask> g1, g2, g3;
	const size_t cnt = sel ? sel->size() : S.size();
	for (size_t i = 0; i < cnt; ++i) {
		const size_t j = sel ? (*sel)[i] : i;
		Site& t = S[j];
		const int* sk = skip ? (*skip)[j] : nullptr;
		GemmDesc d1 = gemm_desc(0, 0, t.a * t.P, t.n * t.c, t.a); d1.skip = sk;
		g1.push_back(gt(d1, t.Lst->data(), v[j], t.T1.get()));
		g2.push_back(gt(stage2(t.a, t.n, t.c, t.P, t.Q, sk), t.Am->data(), t.T1.get(), t.T2.get()));
		GemmDesc d3 = gemm_desc(0, 1, t.a * t.n, t.c, t.Q * t.c); d3.skip = sk;
		g3.push_back(gt(d3, t.T2.get(), t.Rst->data(), y[j]));
		if (rhs) {
			const Lock& L = *t.L;
			const DTensor& Lb = L.rhsL.back();
			const DTensor& Rb = L.rhsR.back();
			const DTensor& bc = L.b->cores[L.curr];
			const size_t beta = Lb.dims[0], beta2 = Rb.dims[0];
			XTT_REQUIRE(Lb.dims[1] == t.a && bc.dims[0] == beta && bc.dims[1] == t.n && bc.dims[2] == beta2 && Rb.dims[1] == t.c,
			            "lockstep ALS: local right-hand side shape mismatch");
			t.Tb.scratch(t.a * t.n * beta2);
			g1.push_back(gt(gemm_desc(1, 0, t.a, t.n * beta2, beta), Lb.data(), bc.data(), t.Tb.get()));
			g2.push_back(gt(gemm_desc(0, 0, t.a * t.n, t.c, beta2), t.Tb.get(), Rb.data(), t.bl.data()));
		}
	}
	gemm_group_f64(g1, s);
	gemm_group_f64(g2, s);
	gemm_group_f64(g3, s);
}

// energy of the SPD functional at each lock's current site: |½ xᵀA_loc x − xᵀb_loc|
void energies(std::vector<Lock*>& Ls, std::vector<double>& E, hipStream_t s) {
	E.assign(Ls.size(), 0.0);
	if (Ls.empty()) return;
	ScratchFrame frame;
	std::vector<Site> S(Ls.size());
	std::vector<DTensor> y(Ls.size());
	std::vector<const double*> v;
	std::vector<double*> yo;
	for (size_t j = 0; j < Ls.size(); ++j) {
		site_init(S[j], Ls[j]);
		y[j] = DTensor::scratch({S[j].a, S[j].n, S[j].c});
		v.push_back(Ls[j]->x->cores[Ls[j]->curr].data());
		yo.push_back(y[j].data());
	}
	local_apply(S, v, yo, true, nullptr, s);
	std::vector<Dot2Task> dt;
	for (size_t j = 0; j < Ls.size(); ++j) {
		const DTensor& X = Ls[j]->x->cores[Ls[j]->curr];
		dt.push_back(Dot2Task{(int)X.size(), X.data(), S[j].bl.data(), X.data(), y[j].data()});
	}
	DBuf<double> out;
	out.scratch(2 * Ls.size());
	dot2_group(dt, out.get(), s);
	std::vector<double> h(2 * Ls.size());
	read_back(h.data(), out.get(), h.size() * sizeof(double), s);
	table_ring_release(s);
	for (size_t j = 0; j < Ls.size(); ++j) E[j] = std::abs(0.5 * h[2 * j + 1] - h[2 * j]);
}

void gemm_plan(const std::vector<GemmTask>& t, GemmGroupPlan& p, hipStream_t s, double) { gemm_group_plan_f64(t, p, s); }
void gemm_plan(const std::vector<GemmTask>& t, GemmGroupPlan& p, hipStream_t s, float) { gemm_group_plan_f32(t, p, s); }
void gemm_run(const GemmGroupPlan& p, hipStream_t s, double) { gemm_group_run_f64(p, s); }
void gemm_run(const GemmGroupPlan& p, hipStream_t s, float) { gemm_group_run_f32(p, s); }

// The three products of q = A_loc·p for system k (its skip flag given), appended to the three launch lists
using ProdFn = std::function<void(size_t k, const int* skip, std::vector<GemmTask>&, std::vector<GemmTask>&, std::vector<GemmTask>&)>;

// CG steps for every system of `tasks` (started already; state k in st[k]) until all are done or maxIter steps ran;
// h receives the final states. The products and the CG step of the systems still iterating keep their tables on
// the device for every iteration of a chunk; the products of a system that converges inside a chunk exit at once
// (skip flag), and the tables are rebuilt without it once the open set has shrunk by a quarter.
template <class T>
void cg_iterate(const std::vector<CGTaskT<T>>& tasks, CGState* st, int maxIter, const ProdFn& prod, std::vector<CGState>& h,
                hipStream_t s) {
	const size_t K = tasks.size();
	GemmGroupPlan g1, g2, g3;
	const bool f64 = sizeof(T) == 8;
	g1.tag = f64 ? "gemm_group_kernel<double> CG L·p" : "gemm_group_kernel<float> CG L·p";
	g2.tag = f64 ? "gemm_group_kernel<double> CG A·T1" : "gemm_group_kernel<float> CG A·T1";
	g3.tag = f64 ? "gemm_group_kernel<double> CG T2·R" : "gemm_group_kernel<float> CG T2·R";
	DBuf<CGTaskT<T>> dTasks;
	DBuf<CGChunk<T>> dChunks;
	DBuf<double> part;
	int planned = 0, chunks = 0;
	// systems above kCGChunk elements are split over several workgroups (the one-workgroup step of a 64·16·64
	// system is bound by one CU's bandwidth); batches of small systems keep the single-launch step
	auto step = [&] {
		if (chunks) cg_chunk_step_run<T>(dChunks.get(), chunks, s);
		else cg_group_step_run<T>(dTasks.get(), planned, s);
	};
	auto plan = [&](const std::vector<size_t>& open) {
		std::vector<GemmTask> t1, t2, t3;
		std::vector<CGTaskT<T>> tk;
		bool split = false;
		for (size_t k : open) {
			prod(k, &st[k].done, t1, t2, t3);
			tk.push_back(tasks[k]);
			split |= tasks[k].n > kCGChunk;
		}
		gemm_plan(t1, g1, s, T());  // each plan also runs once: the first product of the next iteration
		gemm_plan(t2, g2, s, T());
		gemm_plan(t3, g3, s, T());
		planned = (int)open.size();
		chunks = 0;
		if (split) {
			std::vector<CGChunk<T>> ch;
			size_t parts = 0;
			for (const CGTaskT<T>& t : tk) parts += 2 * (size_t)((t.n + kCGChunk - 1) / kCGChunk);
			part.scratch(std::max<size_t>(parts, 1));
			size_t off = 0;
			for (const CGTaskT<T>& t : tk) {
				const int nch = std::max(1, (t.n + kCGChunk - 1) / kCGChunk);
				for (int c = 0; c < nch; ++c)
					ch.push_back(CGChunk<T>{t, c * kCGChunk, std::min(kCGChunk, t.n - c * kCGChunk), nch, c, part.get() + off});
				off += 2 * (size_t)nch;
			}
			dChunks.scratch(ch.size());
			table_upload_to(dChunks.get(), ch.data(), ch.size() * sizeof(CGChunk<T>), s);
			chunks = (int)ch.size();
		} else {
			dTasks.scratch(tk.size());
			table_upload_to(dTasks.get(), tk.data(), tk.size() * sizeof(CGTaskT<T>), s);
		}
		step();
	};
	std::vector<size_t> open(K);
	for (size_t k = 0; k < K; ++k) open[k] = k;
	plan(open);
	h.resize(K);
	int done = 1, chunk = 8;  // iterations run so far; chunks of 8, then 16 (iterations of converged systems are skipped)
	for (;;) {
		read_back(h.data(), st, K * sizeof(CGState), s);
		std::vector<size_t> still;
		for (size_t k = 0; k < K; ++k) if (!h[k].done) still.push_back(k);
		if (still.empty() || done >= maxIter) break;
		int n = std::min(chunk, maxIter - done);
		if (4 * still.size() <= 3 * (size_t)planned) {
			table_ring_release(s);
			plan(still);
			--n;
			done += 1;
		}
		for (int c = 0; c < n; ++c) {
			gemm_run(g1, s, T());
			gemm_run(g2, s, T());
			gemm_run(g3, s, T());
			step();
		}
		done += n;
		chunk = std::min(16, 2 * chunk);
	}
	table_ring_release(s);
}

// states of K systems in the caller's scratch frame: tolerance tol2[k] (one value for all when tol2 has one)
DBuf<CGState> cg_states(size_t K, const std::vector<double>& tol2, int maxIter, hipStream_t s) {
	DBuf<CGState> st;
	st.scratch(K);
	std::vector<CGState> init(K);
	for (size_t k = 0; k < K; ++k) {
		CGState& c = init[k];
		std::memset(&c, 0, sizeof(CGState));
		c.tol2 = tol2.size() == 1 ? tol2[0] : tol2[k];
		c.maxIter = maxIter;
	}
	const void* dInit = table_upload(init.data(), K * sizeof(CGState), s);
	XTT_HIP(hipMemcpyAsync(st.get(), dInit, K * sizeof(CGState), hipMemcpyDeviceToDevice, s));
	return st;
}

// FP64 CG on every lock's local system (warm start from its core, which is updated in place)
void local_solves_f64(std::vector<Lock*>& Ls, const ALSOptions& o, hipStream_t s) {
	const size_t J = Ls.size();
	ScratchFrame frame;  // sites, CG vectors, states and tables
	std::vector<Site> S(J);
	std::vector<DBuf<double>> r(J), p(J), q(J);
	std::vector<CGTask> tasks(J);
	std::vector<const double*> xv(J);
	std::vector<double*> qv(J);
	DBuf<CGState> st = cg_states(J, {o.localTol * o.localTol}, o.maxLocalIter, s);
	for (size_t j = 0; j < J; ++j) {
		site_init(S[j], Ls[j]);
		const size_t N = S[j].a * S[j].n * S[j].c;
		r[j].scratch(N); p[j].scratch(N); q[j].scratch(N);
		double* x = Ls[j]->x->cores[Ls[j]->curr].data();
		tasks[j] = CGTask{(int)N, x, r[j].get(), p[j].get(), q[j].get(), S[j].bl.data(), st.get() + j};
		xv[j] = x; qv[j] = q[j].get();
		Ls[j]->localSolves++;
	}
	local_apply(S, xv, qv, true, nullptr, s);  // q = A_loc·x and the right-hand sides
	cg_group_start(tasks, s);
	const ProdFn prod = [&](size_t j, const int* skip, std::vector<GemmTask>& t1, std::vector<GemmTask>& t2, std::vector<GemmTask>& t3) {
		Site& t = S[j];
		GemmDesc d1 = gemm_desc(0, 0, t.a * t.P, t.n * t.c, t.a); d1.skip = skip;
		t1.push_back(gt(d1, t.Lst->data(), p[j].get(), t.T1.get()));
		t2.push_back(gt(stage2(t.a, t.n, t.c, t.P, t.Q, skip), t.Am->data(), t.T1.get(), t.T2.get()));
		GemmDesc d3 = gemm_desc(0, 1, t.a * t.n, t.c, t.Q * t.c); d3.skip = skip;
		t3.push_back(gt(d3, t.T2.get(), t.Rst->data(), q[j].get()));
	};
	std::vector<CGState> h;
	cg_iterate<double>(tasks, st.get(), o.maxLocalIter, prod, h, s);
	for (size_t j = 0; j < J; ++j) Ls[j]->cgIters += (size_t)std::max(0, h[j].iter);
}

// Mixed precision (als.cpp's cg_mixed, system by system): FP64 iterative refinement around an FP32 CG. Each round
// forms r = b − A·x in FP64 for the systems still open (one grouped residual, one read-back), stops a system when
// ‖r‖ <= tol·‖b‖ or the refinement cap is reached, restarts from x = 0 where the warm start is worse than zero,
// solves A·e = r/‖r‖ from e = 0 in FP32 to the inner tolerance (inner_tolerance: no tighter than the outer test
// needs) and adds ‖r‖·e to x in FP64.
void local_solves_mixed(std::vector<Lock*>& Ls, const ALSOptions& o, hipStream_t s) {
	const size_t J = Ls.size();
	const double inner = o.innerTol > 0 ? o.innerTol : 1e-5, tol2 = o.localTol * o.localTol;
	ScratchFrame frame;  // sites, FP32 operator copies, vectors, states and tables
	struct Low {
		DBuf<float> L, Am, R, rhs, e, r, p, q, T1, T2;
		bool staged = false;
	};
	std::vector<Site> S(J);
	std::vector<Low> lw(J);
	std::vector<DBuf<double>> r(J), q(J);
	std::vector<const double*> xv(J);
	std::vector<double*> qv(J);
	std::vector<size_t> N(J), open(J);
	for (size_t j = 0; j < J; ++j) {
		site_init(S[j], Ls[j]);
		N[j] = S[j].a * S[j].n * S[j].c;
		r[j].scratch(N[j]); q[j].scratch(N[j]);
		xv[j] = Ls[j]->x->cores[Ls[j]->curr].data();
		qv[j] = q[j].get();
		open[j] = j;
		Ls[j]->localSolves++;
	}
	local_apply(S, xv, qv, true, nullptr, s);  // q = A_loc·x and the right-hand sides
	DBuf<double> norms;
	norms.scratch(2 * J);
	std::vector<double> hn(2 * J);
	for (int ref = 0;; ++ref) {
		std::vector<ResidTask> rt;
		for (size_t j : open) rt.push_back(ResidTask{(int)N[j], r[j].get(), S[j].bl.data(), q[j].get()});
		resid_group(rt, norms.get(), s);
		read_back(hn.data(), norms.get(), 2 * open.size() * sizeof(double), s);
		table_ring_release(s);
		std::vector<size_t> next;
		std::vector<double> rn, innerTol2;
		std::vector<MixTask> mix;
		for (size_t i = 0; i < open.size(); ++i) {
			const size_t j = open[i];
			double rr = hn[2 * i];
			const double bb = hn[2 * i + 1];
			if (rr <= tol2 * bb || ref == o.maxRefine) continue;
			double* x = Ls[j]->x->cores[Ls[j]->curr].data();
			const double* src = r[j].get();
			if (ref == 0 && rr > bb) {  // the warm start is worse than zero: refine from x = 0 (r = b)
				mix.push_back(MixTask{MIX_ZERO, (long long)N[j], S[j].bl.data(), x, r[j].get(), 0.0});
				src = S[j].bl.data();
				rr = bb;
			}
			Low& w = lw[j];
			Site& t = S[j];
			if (!w.staged) {
				w.L.scratch(t.Lst->size()); w.Am.scratch(t.Am->size()); w.R.scratch(t.Rst->size());
				w.rhs.scratch(N[j]); w.e.scratch(N[j]); w.r.scratch(N[j]); w.p.scratch(N[j]); w.q.scratch(N[j]);
				w.T1.scratch(t.a * t.P * t.n * t.c); w.T2.scratch(t.a * t.n * t.Q * t.c);
				mix.push_back(MixTask{MIX_CVT32, (long long)t.Lst->size(), t.Lst->data(), w.L.get(), nullptr, 1.0});
				mix.push_back(MixTask{MIX_CVT32, (long long)t.Am->size(), t.Am->data(), w.Am.get(), nullptr, 1.0});
				mix.push_back(MixTask{MIX_CVT32, (long long)t.Rst->size(), t.Rst->data(), w.R.get(), nullptr, 1.0});
				w.staged = true;
			}
			mix.push_back(MixTask{MIX_CVT32, (long long)N[j], src, w.rhs.get(), nullptr, 1.0 / std::sqrt(rr)});
			next.push_back(j);
			rn.push_back(std::sqrt(rr));
			const double it = inner_tolerance(inner, tol2 * bb, rr);
			innerTol2.push_back(it * it);
		}
		if (next.empty()) break;
		mix_group(mix, s);
		const size_t K = next.size();
		DBuf<CGState> st = cg_states(K, innerTol2, o.maxLocalIter, s);
		std::vector<CGTaskT<float>> tasks(K), start(K);
		for (size_t k = 0; k < K; ++k) {
			Low& w = lw[next[k]];
			tasks[k] = CGTaskT<float>{(int)N[next[k]], w.e.get(), w.r.get(), w.p.get(), w.q.get(), w.rhs.get(), st.get() + k};
			start[k] = tasks[k];
			start[k].q = nullptr;  // e = 0
			Ls[next[k]]->refinements++;
		}
		cg_group_start(start, s);
		const ProdFn prod = [&](size_t k, const int* skip, std::vector<GemmTask>& t1, std::vector<GemmTask>& t2, std::vector<GemmTask>& t3) {
			Site& t = S[next[k]];
			Low& w = lw[next[k]];
			GemmDesc d1 = gemm_desc(0, 0, t.a * t.P, t.n * t.c, t.a); d1.skip = skip;
			t1.push_back(GemmTask{d1, w.L.get(), w.p.get(), w.T1.get()});
			t2.push_back(GemmTask{stage2(t.a, t.n, t.c, t.P, t.Q, skip), w.Am.get(), w.T1.get(), w.T2.get()});
			GemmDesc d3 = gemm_desc(0, 1, t.a * t.n, t.c, t.Q * t.c); d3.skip = skip;
			t3.push_back(GemmTask{d3, w.T2.get(), w.R.get(), w.q.get()});
		};
		std::vector<CGState> h;
		cg_iterate<float>(tasks, st.get(), o.maxLocalIter, prod, h, s);
		mix.clear();
		for (size_t k = 0; k < K; ++k) {
			const size_t j = next[k];
			Ls[j]->cgIters += (size_t)std::max(0, h[k].iter);
			mix.push_back(MixTask{MIX_AXPY32, (long long)N[j], lw[j].e.get(), Ls[j]->x->cores[Ls[j]->curr].data(), nullptr, rn[k]});
		}
		mix_group(mix, s);  // x += ‖r‖·e
		local_apply(S, xv, qv, false, nullptr, s, &next);  // q = A_loc·x
		open = std::move(next);
	}
	table_ring_release(s);
}

void local_solves(std::vector<Lock*>& Ls, const ALSOptions& o, hipStream_t s) {
	if (Ls.empty()) return;
	if (o.precision == P_F64) local_solves_f64(Ls, o, s); else local_solves_mixed(Ls, o, s);
}

// ---- core moves: QR to the right (x_k = Q, x_{k+1} = R·x_{k+1}), RQ to the left (x_k = Q, x_{k-1} = x_{k-1}·R) --
struct Move {
	TT* x;
	size_t k;
	bool right;
};
void moves(const std::vector<Move>& M, hipStream_t s) {
	if (M.empty()) return;
	const size_t J = M.size();
	ScratchFrame frame;  // B, Qt, Rt and R; the new cores Q and T are owned
	std::vector<DTensor> B(J), Qt(J), Rt(J), Q(J), R(J);
	std::vector<QRJob> jobs;
	std::vector<MapTask> maps;
	for (size_t j = 0; j < J; ++j) {
		const DTensor& F = M[j].x->cores[M[j].k];
		const size_t a = F.dims[0], n = F.dims[1], c = F.dims[2];
		if (M[j].right) {
			const size_t m = a * n, kk = std::min(m, c);
			Q[j] = DTensor({a, n, kk});
			R[j] = DTensor::scratch({kk, c});
			jobs.push_back(QRJob{F.data(), (int)m, (int)c, false, Q[j].data(), R[j].data()});
		} else {  // QR(J·Fᵀ·J) = Q̃R̃ ⇒ F = (J·R̃ᵀ·J)(J·Q̃ᵀ·J)
			const size_t m = a, nn = n * c, kk = std::min(m, nn);
			B[j] = DTensor::scratch({nn, m});
			maps.push_back(map_revtrans(F.data(), m, nn, B[j].data()));
			Qt[j] = DTensor::scratch({nn, kk});
			Rt[j] = DTensor::scratch({kk, m});
			jobs.push_back(QRJob{B[j].data(), (int)nn, (int)m, false, Qt[j].data(), Rt[j].data()});
		}
	}
	map_group(maps, s);
	hh_qr_group(jobs, s);
	maps.clear();
	for (size_t j = 0; j < J; ++j) {
		if (M[j].right) continue;
		const DTensor& F = M[j].x->cores[M[j].k];
		const size_t a = F.dims[0], n = F.dims[1], c = F.dims[2], nn = n * c, kk = std::min(a, nn);
		R[j] = DTensor::scratch({a, kk});
		Q[j] = DTensor({kk, n, c});
		maps.push_back(map_revtrans(Rt[j].data(), kk, a, R[j].data()));
		maps.push_back(map_revtrans(Qt[j].data(), nn, kk, Q[j].data()));
	}
	map_group(maps, s);
	std::vector<GemmTask> g;
	std::vector<DTensor> T(J);
	for (size_t j = 0; j < J; ++j) {
		TT& x = *M[j].x;
		if (M[j].right) {
			const DTensor& N = x.cores[M[j].k + 1];
			const size_t kk = R[j].dims[0], c = R[j].dims[1];
			T[j] = DTensor({kk, N.dims[1], N.dims[2]});
			g.push_back(gt(gemm_desc(0, 0, kk, N.dims[1] * N.dims[2], c), R[j].data(), N.data(), T[j].data()));
		} else {
			const DTensor& N = x.cores[M[j].k - 1];
			const size_t a = R[j].dims[0], kk = R[j].dims[1];
			T[j] = DTensor({N.dims[0], N.dims[1], kk});
			g.push_back(gt(gemm_desc(0, 0, N.dims[0] * N.dims[1], kk, a), N.data(), R[j].data(), T[j].data()));
		}
	}
	gemm_group_f64(g, s);
	for (size_t j = 0; j < J; ++j) {
		TT& x = *M[j].x;
		const size_t k = M[j].k;
		x.cores[k] = std::move(Q[j]);
		x.cores[M[j].right ? k + 1 : k - 1] = std::move(T[j]);
		x.corePosition = M[j].right ? k + 1 : k - 1;
		x.canonicalized = true;
	}
}

// ---- one ALS_SPD call (sites = 1) for every problem --------------------------------------------------------------
void als_lockstep(const ALSOptions& o, std::vector<Lock>& P, hipStream_t s) {
	for (Lock& L : P) {
		L.canonAtEnd = L.x->canonicalized;
		L.corePosAtEnd = L.x->corePosition;
		L.opL.clear(); L.opR.clear(); L.rhsL.clear(); L.rhsR.clear(); L.am.clear(); L.amM.clear();
		L.le2 = 1e102; L.le = 1e101; L.en = 1e100; L.halfSweeps = 0; L.increasing = true; L.active = true;
	}
	prepare_x_all(P, s);
	// permuted operator cores and the stacks' boundary entries
	std::vector<MapTask> maps;
	for (Lock& L : P) {
		for (size_t k = 0; k < L.d; ++k) {
			const DTensor& Ac = L.A->cores[k];
			DTensor a({Ac.dims[1], Ac.dims[3], Ac.dims[0], Ac.dims[2]}), m({Ac.dims[1], Ac.dims[0], Ac.dims[3], Ac.dims[2]});
			maps.push_back(map_permute(4, Ac.dims.data(), kAm.data(), Ac.data(), a.data()));
			maps.push_back(map_permute(4, Ac.dims.data(), kAmM.data(), Ac.data(), m.data()));
			L.am.push_back(std::move(a));
			L.amM.push_back(std::move(m));
		}
		L.opL.push_back(ones_dev({1, 1, 1}, maps));
		L.opR.push_back(ones_dev({1, 1, 1}, maps));
		L.rhsL.push_back(ones_dev({1, 1}, maps));
		L.rhsR.push_back(ones_dev({1, 1}, maps));
	}
	map_group(maps, s);
	// prepare_stacks: right pushes from the last core down to first + 1, then left pushes below first
	for (size_t step = 0;; ++step) {
		std::vector<Push> pu;
		for (Lock& L : P) {
			if (L.d < 2 + step) continue;
			const size_t i = L.d - 1 - step;
			if (i > L.first) pu.push_back(Push{&L, i, false});
		}
		if (pu.empty()) break;
		pushes(pu, s);
	}
	for (size_t i = 0;; ++i) {
		std::vector<Push> pu;
		for (Lock& L : P) if (i < L.first) pu.push_back(Push{&L, i, true});
		if (pu.empty()) break;
		pushes(pu, s);
	}
	// the start energy at the first site
	std::vector<Lock*> act;
	for (Lock& L : P) { L.curr = L.first; act.push_back(&L); }
	std::vector<double> E;
	energies(act, E, s);
	for (size_t j = 0; j < act.size(); ++j) act[j]->en = E[j];
	// sweeps: local solve, end-of-sweep test, move to the next site
	while (!act.empty()) {
		local_solves(act, o, s);
		std::vector<Lock*> atEnd;
		for (Lock* L : act)
			if ((!L->increasing && L->curr == L->first) || (L->increasing && L->curr == L->second - 1)) atEnd.push_back(L);
		energies(atEnd, E, s);
		std::vector<Lock*> stay;
		std::vector<size_t> endIdx(act.size(), SIZE_MAX);
		for (size_t j = 0; j < atEnd.size(); ++j) {
			Lock& L = *atEnd[j];
			L.halfSweeps += 1;
			L.halfSweepsAll += 1;
			L.le2 = L.le;
			L.le = L.en;
			L.en = E[j];
			const bool stop = L.halfSweeps == o.numHalfSweeps || std::abs(L.le - L.en) < o.convergenceEpsilon ||
			                  std::abs(L.le2 - L.en) < o.convergenceEpsilon || L.second - L.first <= 1;
			if (stop) L.active = false;
			else L.increasing = !L.increasing;
		}
		std::vector<Move> mv;
		std::vector<Push> pu;
		std::vector<Lock*> moving;
		for (Lock* L : act) {
			if (!L->active) continue;
			mv.push_back(Move{L->x, L->curr, L->increasing});
			moving.push_back(L);
		}
		moves(mv, s);
		for (Lock* L : moving) {
			if (L->increasing) {
				L->opR.pop_back(); L->rhsR.pop_back();
				pu.push_back(Push{L, L->curr, true});
			} else {
				L->opL.pop_back(); L->rhsL.pop_back();
				pu.push_back(Push{L, L->curr, false});
			}
		}
		pushes(pu, s);
		for (Lock* L : moving) L->curr += L->increasing ? 1 : -1;
		act = moving;
	}
	// preserveCorePosition: back to the entry position (keeping ranks)
	for (;;) {
		std::vector<Move> mv;
		for (Lock& L : P) {
			TT& x = *L.x;
			if (!(o.preserveCorePosition && L.canonAtEnd) || x.corePosition == L.corePosAtEnd) continue;
			const bool right = x.corePosition < L.corePosAtEnd;
			mv.push_back(Move{&x, x.corePosition, right});
		}
		if (mv.empty()) break;
		moves(mv, s);
	}
	for (Lock& L : P) { L.opL.clear(); L.opR.clear(); L.rhsL.clear(); L.rhsR.clear(); L.am.clear(); L.amM.clear(); }
}

// ---- round(eps) (ttNetwork.cpp:643-665) for every problem ------------------------------------------------------
bool exceeds_max(const TT& x) {
	for (size_t i = 0; i < x.order(); ++i) {
		const DTensor& c = x.cores[i];
		const size_t e = x.ext(i);
		if (c.dims.front() > e * c.dims.back() || c.dims.back() > e * c.dims.front()) return true;
	}
	return false;
}

// QC moves k → k+1 (rank reduction allowed) for the listed problems
void qc_moves(const std::vector<std::pair<TT*, size_t>>& M, hipStream_t s) {
	const size_t J = M.size();
	if (!J) return;
	ScratchFrame frame;  // C, pivots and ranks; Q becomes the core (owned)
	std::vector<DBuf<double>> q(J), c(J);
	std::vector<DBuf<int>> jp(J);
	DBuf<int> rk;
	rk.scratch(J);
	std::vector<QCJob> jobs;
	for (size_t j = 0; j < J; ++j) {
		const DTensor& F = M[j].first->cores[M[j].second];
		const size_t m = F.dims[0] * F.dims[1], n = F.dims[2], k = std::min(m, n);
		q[j].resize(m * k); c[j].scratch(k * n); jp[j].scratch(n);
		jobs.push_back(QCJob{F.data(), (int)m, (int)n, q[j].get(), c[j].get(), jp[j].get(), rk.get() + j});
	}
	hh_qc_group(jobs, s);
	std::vector<int> h(J);
	read_back(h.data(), rk.get(), J * sizeof(int), s);
	table_ring_release(s);
	std::vector<GemmTask> g;
	std::vector<DTensor> T(J);
	for (size_t j = 0; j < J; ++j) {
		TT& x = *M[j].first;
		const size_t k = M[j].second, r = (size_t)h[j];
		const DTensor& F = x.cores[k];
		const DTensor& N = x.cores[k + 1];
		T[j] = DTensor({r, N.dims[1], N.dims[2]});
		g.push_back(gt(gemm_desc(0, 0, r, N.dims[1] * N.dims[2], F.dims[2]), c[j].get(), N.data(), T[j].data()));
	}
	gemm_group_f64(g, s);
	for (size_t j = 0; j < J; ++j) {
		TT& x = *M[j].first;
		const size_t k = M[j].second, r = (size_t)h[j];
		const DTensor& F = x.cores[k];
		x.cores[k] = DTensor({F.dims[0], F.dims[1], r}, std::move(q[j]));
		x.cores[k + 1] = std::move(T[j]);
		x.corePosition = k + 1;
		x.canonicalized = true;
	}
}

// CQ moves k → k−1 (rank reduction allowed): F (r × n·r') = C·Q by a pivoted QR of Fᵀ; x_k = Q, x_{k-1} = x_{k-1}·C
void cq_moves(const std::vector<std::pair<TT*, size_t>>& M, hipStream_t s) {
	const size_t J = M.size();
	if (!J) return;
	ScratchFrame frame;  // Fᵀ, its factors, C, pivots and ranks; the new cores are owned
	std::vector<DTensor> Ft(J);
	std::vector<DBuf<double>> q(J), c(J);
	std::vector<DBuf<int>> jp(J);
	DBuf<int> rk;
	rk.scratch(J);
	std::vector<MapTask> maps;
	std::vector<QCJob> jobs;
	for (size_t j = 0; j < J; ++j) {
		const DTensor& F = M[j].first->cores[M[j].second];
		const size_t r = F.dims[0], nr = F.dims[1] * F.dims[2], k = std::min(nr, r);
		Ft[j] = DTensor::scratch({nr, r});
		const size_t d2[2] = {r, nr}, sh[2] = {1, 0};
		maps.push_back(map_permute(2, d2, sh, F.data(), Ft[j].data()));
		q[j].scratch(nr * k); c[j].scratch(k * r); jp[j].scratch(r);
		jobs.push_back(QCJob{Ft[j].data(), (int)nr, (int)r, q[j].get(), c[j].get(), jp[j].get(), rk.get() + j});
	}
	map_group(maps, s);
	hh_qc_group(jobs, s);
	std::vector<int> hr(J);
	read_back(hr.data(), rk.get(), J * sizeof(int), s);
	table_ring_release(s);
	std::vector<DTensor> C(J), Q(J), T(J);
	maps.clear();
	std::vector<GemmTask> g;
	for (size_t j = 0; j < J; ++j) {
		TT& x = *M[j].first;
		const DTensor& F = x.cores[M[j].second];
		const DTensor& N = x.cores[M[j].second - 1];
		const size_t r = F.dims[0], nr = F.dims[1] * F.dims[2], rr = (size_t)hr[j];
		C[j] = DTensor::scratch({r, rr});
		Q[j] = DTensor({rr, F.dims[1], F.dims[2]});
		{ const size_t d2[2] = {rr, r}, sh[2] = {1, 0}; maps.push_back(map_permute(2, d2, sh, c[j].get(), C[j].data())); }
		{ const size_t d2[2] = {nr, rr}, sh[2] = {1, 0}; maps.push_back(map_permute(2, d2, sh, q[j].get(), Q[j].data())); }
		T[j] = DTensor({N.dims[0], N.dims[1], rr});
		g.push_back(gt(gemm_desc(0, 0, N.dims[0] * N.dims[1], rr, r), N.data(), C[j].data(), T[j].data()));
	}
	map_group(maps, s);
	gemm_group_f64(g, s);
	for (size_t j = 0; j < J; ++j) {
		TT& x = *M[j].first;
		const size_t k = M[j].second;
		x.cores[k] = std::move(Q[j]);
		x.cores[k - 1] = std::move(T[j]);
		x.corePosition = k - 1;
		x.canonicalized = true;
	}
}

// move_core's sweeps while a core exceeds its maximal ranks (ttNetwork.cpp:609-622), in lockstep, for problems whose
// core sits at `pos` = their last core: to the first core (CQ moves), to the last (QC moves), until no core exceeds
void max_rank_sweeps(std::vector<Lock*> E, hipStream_t s) {
	while (!E.empty()) {
		for (;;) {
			std::vector<std::pair<TT*, size_t>> M;
			for (Lock* L : E) if (L->x->corePosition > 0) M.push_back({L->x, L->x->corePosition});
			if (M.empty()) break;
			cq_moves(M, s);
		}
		for (;;) {
			std::vector<std::pair<TT*, size_t>> M;
			for (Lock* L : E) if (L->x->corePosition + 1 < L->d) M.push_back({L->x, L->x->corePosition});
			if (M.empty()) break;
			qc_moves(M, s);
		}
		std::vector<Lock*> still;
		for (Lock* L : E) if (exceeds_max(*L->x)) still.push_back(L);
		E.swap(still);
	}
}

// round_edge(k, k−1) with the left part orthogonal (tt.cpp round_edge, first branch) for the listed problems
void edges(const std::vector<std::pair<TT*, size_t>>& M, double eps, hipStream_t s) {
	const size_t J = M.size();
	if (!J) return;
	ScratchFrame frame;  // everything but the new cores Fn and Tn
	// CQ of F (r × n·r'): pivoted QR of Fᵀ
	std::vector<DTensor> Ft(J);
	std::vector<DBuf<double>> q(J), c(J);
	std::vector<DBuf<int>> jp(J);
	DBuf<int> rk;
	rk.scratch(J);
	std::vector<MapTask> maps;
	std::vector<QCJob> jobs;
	for (size_t j = 0; j < J; ++j) {
		const DTensor& F = M[j].first->cores[M[j].second];
		const size_t r = F.dims[0], nr = F.dims[1] * F.dims[2], k = std::min(nr, r);
		Ft[j] = DTensor::scratch({nr, r});
		const size_t d2[2] = {r, nr}, sh[2] = {1, 0};
		maps.push_back(map_permute(2, d2, sh, F.data(), Ft[j].data()));
		q[j].scratch(nr * k); c[j].scratch(k * r); jp[j].scratch(r);
		jobs.push_back(QCJob{Ft[j].data(), (int)nr, (int)r, q[j].get(), c[j].get(), jp[j].get(), rk.get() + j});
	}
	map_group(maps, s);
	hh_qc_group(jobs, s);
	std::vector<int> hr(J);
	read_back(hr.data(), rk.get(), J * sizeof(int), s);
	table_ring_release(s);
	// coreA = Ctᵀ (r × rk), Fq = Qtᵀ (rk × n·r'); SVD of X = coreAᵀ (rk × r) by Jacobi on coreA (transposed output)
	std::vector<DTensor> coreA(J), Fq(J);
	std::vector<DBuf<double>> u(J), sv(J), vt(J), ws(J);
	std::vector<SVDJob> sj;
	maps.clear();
	for (size_t j = 0; j < J; ++j) {
		const DTensor& F = M[j].first->cores[M[j].second];
		const size_t r = F.dims[0], nr = F.dims[1] * F.dims[2], rk = (size_t)hr[j];
		coreA[j] = DTensor::scratch({r, rk});
		Fq[j] = DTensor::scratch({rk, F.dims[1], F.dims[2]});
		{ const size_t d2[2] = {rk, r}, sh[2] = {1, 0}; maps.push_back(map_permute(2, d2, sh, c[j].get(), coreA[j].data())); }
		{ const size_t d2[2] = {nr, rk}, sh[2] = {1, 0}; maps.push_back(map_permute(2, d2, sh, q[j].get(), Fq[j].data())); }
		// X is rk × r with rk <= r: U = V' (rk × rk) and Vt = U'ᵀ (rk × r)
		const size_t m = rk, n = r, k = std::min(m, n);
		u[j].scratch(m * k); sv[j].scratch(k); vt[j].scratch(k * n);
		const size_t M2 = std::max(m, n), N2 = k, np = N2 + (N2 & 1);
		ws[j].scratch(M2 * np + np * np + 2 * np);
		sj.push_back(SVDJob{coreA[j].data(), (int)r, (int)rk, vt[j].get(), sv[j].get(), u[j].get(), ws[j].get(), true});
	}
	map_group(maps, s);
	jacobi_svd_group(sj, s);
	// singular values → truncation ranks (xerus: min(k, maxRank) then the first j >= 1 with s_j <= eps·s_0)
	std::vector<size_t> off(J + 1, 0);
	for (size_t j = 0; j < J; ++j) off[j + 1] = off[j] + sv[j].size();
	std::vector<double> S(off[J]);
	{
		DBuf<double> all;
		all.scratch(off[J]);
		maps.clear();
		for (size_t j = 0; j < J; ++j) maps.push_back(map_copy(sv[j].get(), all.get() + off[j], sv[j].size()));
		map_group(maps, s);
		read_back(S.data(), all.get(), off[J] * sizeof(double), s);
		table_ring_release(s);
	}
	std::vector<size_t> rank(J);
	for (size_t j = 0; j < J; ++j) {
		const double* sh = S.data() + off[j];
		const size_t k = sv[j].size();
		if (std::isnan(sh[0])) throw Error(ERR_NUMERIC, "SVD: the one-sided Jacobi iteration did not converge in 60 sweeps");
		size_t rr = k;
		for (size_t i = 1; i < rr; ++i) if (sh[i] <= eps * sh[0]) { rr = i; break; }
		rank[j] = rr;
	}
	// U (rk × rank) = leading columns of u (read in place), Vt (rank × r) = leading rows of vt, scaled by S (a
	// second launch: the tasks of one launch run concurrently)
	std::vector<DTensor> Vt(J), Fn(J), Tn(J);
	std::vector<MapTask> scale;
	maps.clear();
	std::vector<GemmTask> g;
	for (size_t j = 0; j < J; ++j) {
		TT& x = *M[j].first;
		const size_t k0 = M[j].second;
		const DTensor& F = x.cores[k0];
		const DTensor& T = x.cores[k0 - 1];
		const size_t r = F.dims[0], rk = (size_t)hr[j], kk = sv[j].size(), rr = rank[j];
		Vt[j] = DTensor::scratch({rr, r});
		maps.push_back(map_copy(vt[j].get(), Vt[j].data(), rr * r));
		scale.push_back(map_scale_by(Vt[j].data(), sv[j].get(), rr, r, true));
		Fn[j] = DTensor({rr, F.dims[1], F.dims[2]});
		Tn[j] = DTensor({T.dims[0], T.dims[1], rr});
		GemmDesc du = gemm_desc(1, 0, rr, F.dims[1] * F.dims[2], rk);
		du.lda = kk;  // U = u[:, :rr] read in place (ld kk)
		g.push_back(gt(du, u[j].get(), Fq[j].data(), Fn[j].data()));
	}
	map_group(maps, s);
	map_group(scale, s);
	std::vector<GemmTask> g2;
	for (size_t j = 0; j < J; ++j) {
		TT& x = *M[j].first;
		const DTensor& T = x.cores[M[j].second - 1];
		const size_t r = x.cores[M[j].second].dims[0], rr = rank[j];
		g2.push_back(gt(gemm_desc(0, 1, T.dims[0] * T.dims[1], rr, r), T.data(), Vt[j].data(), Tn[j].data()));
	}
	gemm_group_f64(g, s);
	gemm_group_f64(g2, s);
	for (size_t j = 0; j < J; ++j) {
		TT& x = *M[j].first;
		const size_t k0 = M[j].second;
		x.cores[k0] = std::move(Fn[j]);
		x.cores[k0 - 1] = std::move(Tn[j]);
	}
}

// round_edge(k, k−1)'s second branch (5|F||T| < 6r⁴: the SVD of the merged pair) for the listed problems:
// X = Fᵀ·Tᵀ (n·r' × r''·n'), its Jacobi SVD (on Xᵀ when X is wider than tall, as calculate_svd does), the truncation
// rank from the singular values, F = Uᵀ's leading rows (rank, n, r'), T = (S·Vt)ᵀ (r'', n', rank)
void edges2(const std::vector<std::pair<TT*, size_t>>& M, double eps, hipStream_t s) {
	const size_t J = M.size();
	if (!J) return;
	ScratchFrame frame;  // X, Xᵀ, the factors and the workspaces; the new cores are owned
	std::vector<DTensor> X(J), Xt(J);
	std::vector<DBuf<double>> u(J), sv(J), vt(J), ws(J);
	std::vector<size_t> mm(J), nn(J), kk(J);
	std::vector<GemmTask> g;
	for (size_t j = 0; j < J; ++j) {
		const DTensor& F = M[j].first->cores[M[j].second];
		const DTensor& T = M[j].first->cores[M[j].second - 1];
		const size_t r = F.dims[0], m = F.dims[1] * F.dims[2], n = T.dims[0] * T.dims[1], k = std::min(m, n);
		XTT_REQUIRE(T.dims[2] == r, "lockstep round: rank mismatch");
		mm[j] = m; nn[j] = n; kk[j] = k;
		X[j] = DTensor::scratch({m, n});
		g.push_back(gt(gemm_desc(1, 1, m, n, r), F.data(), T.data(), X[j].data()));
		u[j].scratch(m * k); sv[j].scratch(k); vt[j].scratch(k * n);
		const size_t np = k + (k & 1);
		ws[j].scratch(std::max(m, n) * np + np * np + 2 * np);
	}
	gemm_group_f64(g, s);
	std::vector<MapTask> maps;
	std::vector<SVDJob> sj;
	for (size_t j = 0; j < J; ++j) {
		const size_t m = mm[j], n = nn[j];
		if (m >= n) {
			sj.push_back(SVDJob{X[j].data(), (int)m, (int)n, u[j].get(), sv[j].get(), vt[j].get(), ws[j].get(), false});
		} else {  // Xᵀ = U'·S·V'ᵀ ⇒ X = V'·S·U'ᵀ, written transposed: U = V' (m × m), Vt = U'ᵀ (m × n)
			Xt[j] = DTensor::scratch({n, m});
			const size_t d2[2] = {m, n}, sh[2] = {1, 0};
			maps.push_back(map_permute(2, d2, sh, X[j].data(), Xt[j].data()));
			sj.push_back(SVDJob{Xt[j].data(), (int)n, (int)m, vt[j].get(), sv[j].get(), u[j].get(), ws[j].get(), true});
		}
	}
	map_group(maps, s);
	jacobi_svd_group(sj, s);
	std::vector<size_t> off(J + 1, 0);
	for (size_t j = 0; j < J; ++j) off[j + 1] = off[j] + kk[j];
	std::vector<double> S(off[J]);
	{
		DBuf<double> all;
		all.scratch(off[J]);
		maps.clear();
		for (size_t j = 0; j < J; ++j) maps.push_back(map_copy(sv[j].get(), all.get() + off[j], kk[j]));
		map_group(maps, s);
		read_back(S.data(), all.get(), off[J] * sizeof(double), s);
		table_ring_release(s);
	}
	std::vector<size_t> rank(J);
	for (size_t j = 0; j < J; ++j) {
		const double* sh = S.data() + off[j];
		if (std::isnan(sh[0])) throw Error(ERR_NUMERIC, "SVD: the one-sided Jacobi iteration did not converge in 60 sweeps");
		size_t rr = kk[j];
		for (size_t i = 1; i < rr; ++i) if (sh[i] <= eps * sh[0]) { rr = i; break; }
		rank[j] = rr;
	}
	// F = the leading rank rows of uᵀ (k × m), T = (the leading rank rows of vt, scaled by S)ᵀ: the scaling is its
	// own launch (the tasks of one launch run concurrently)
	std::vector<DBuf<double>> fo(J);
	maps.clear();
	std::vector<MapTask> scale, tr;
	std::vector<DTensor> Tn(J);
	for (size_t j = 0; j < J; ++j) {
		const size_t m = mm[j], n = nn[j], k = kk[j], rr = rank[j];
		fo[j].resize(k * m);
		const size_t du[2] = {m, k}, sh[2] = {1, 0};
		maps.push_back(map_permute(2, du, sh, u[j].get(), fo[j].get()));
		scale.push_back(map_scale_by(vt[j].get(), sv[j].get(), rr, n, true));
		const DTensor& T = M[j].first->cores[M[j].second - 1];
		Tn[j] = DTensor({T.dims[0], T.dims[1], rr});
		const size_t dv[2] = {rr, n};
		tr.push_back(map_permute(2, dv, sh, vt[j].get(), Tn[j].data()));
	}
	map_group(maps, s);
	map_group(scale, s);
	map_group(tr, s);
	for (size_t j = 0; j < J; ++j) {
		TT& x = *M[j].first;
		const size_t k0 = M[j].second;
		const DTensor& F = x.cores[k0];
		const dims_t fd{rank[j], F.dims[1], F.dims[2]};
		x.cores[k0] = DTensor(fd, std::move(fo[j]));
		x.cores[k0 - 1] = std::move(Tn[j]);
	}
}

void round_lockstep(std::vector<Lock>& P, double eps, hipStream_t s) {
	std::vector<char> initCanon(P.size());
	std::vector<size_t> initPos(P.size());
	for (size_t j = 0; j < P.size(); ++j) {
		TT& x = *P[j].x;
		initCanon[j] = x.canonicalized;
		initPos[j] = x.corePosition;
		XTT_REQUIRE(x.canonicalized, "lockstep round: x must be canonical");
	}
	// move_core(x, d − 1, false): QC moves from the core position to the last core
	for (;;) {
		std::vector<std::pair<TT*, size_t>> M;
		for (Lock& L : P) if (L.x->corePosition + 1 < L.d) M.push_back({L.x, L.x->corePosition});
		if (M.empty()) break;
		qc_moves(M, s);
	}
	// a rank dropped by a QC can leave a neighbour above its maximal rank; move_core then sweeps until it is not
	// (ttNetwork.cpp:609-622)
	{
		std::vector<Lock*> E;
		for (Lock& L : P) if (exceeds_max(*L.x)) E.push_back(&L);
		max_rank_sweeps(E, s);
	}
	// round_edge(d−1−i, d−2−i) from the right, each edge in the branch round_edge takes for it
	for (size_t i = 0;; ++i) {
		std::vector<std::pair<TT*, size_t>> M, M2;
		for (Lock& L : P) {
			if (i + 1 >= L.d) continue;
			TT& x = *L.x;
			const size_t from = L.d - 1 - i;
			const DTensor& F = x.cores[from];
			const DTensor& T = x.cores[from - 1];
			const double r = (double)F.dims[0];
			if (5.0 * (double)F.size() * (double)T.size() >= 6.0 * r * r * r * r) M.push_back({&x, from});
			else M2.push_back({&x, from});
		}
		bool any = false;
		for (Lock& L : P) any |= i + 1 < L.d;
		if (!any) break;
		edges(M, eps, s);
		edges2(M2, eps, s);
	}
	for (size_t j = 0; j < P.size(); ++j) {
		TT& x = *P[j].x;
		x.corePosition = 0;
		x.canonicalized = true;
		if (initCanon[j] && initPos[j] != 0) move_core(x, initPos[j], false);
	}
}

}  // namespace

// ---- ‖A x − b‖ / ‖b‖ per problem (tt_residual): w = A·x + (−b) as a TT, an R-only QR sweep from the left
// (c_0 = w_0, c_k = R_{k-1}·w_k, R_k = R of c_k) and ‖c_{d-1}‖; ‖b‖² from the transfer matrices of b (tt_dot).
// One step per core index for all problems: the cores of w are assembled, multiplied and factored in grouped
// launches, and nothing is read back before the end.
void residuals_lockstep(std::vector<Lock>& P, double* out, hipStream_t s) {
	const size_t J = P.size();
	size_t maxd = 0;
	for (const Lock& L : P) maxd = std::max(maxd, L.d);
	std::vector<DTensor> Rk(J), G(J), last(J);
	DBuf<double> gfin(J);
	for (size_t k = 0; k < maxd; ++k) {
		ScratchFrame frame;  // this step's cores of w, the products M and the factorisations' work buffers
		std::vector<ResCoreTask> asmT;
		std::vector<DTensor> W(J), C(J), M(J), Gn(J);
		for (size_t j = 0; j < J; ++j) {
			const Lock& L = P[j];
			if (k >= L.d) continue;
			const DTensor &Ac = L.A->cores[k], &Xc = L.x->cores[k], &Bc = L.b->cores[k];
			ResCoreTask t{Ac.data(), Xc.data(), Bc.data(), nullptr, (int)Ac.dims[0], (int)Ac.dims[3], (int)Xc.dims[0], (int)Xc.dims[2],
			              (int)Bc.dims[0], (int)Bc.dims[2], (int)Ac.dims[1], (int)Ac.dims[2], k == 0, k + 1 == L.d};
			XTT_REQUIRE(Ac.dims[2] == Xc.dims[1] && Bc.dims[1] == Ac.dims[1], "lockstep residual: dimension mismatch");
			const size_t Lw = t.first ? 1 : (size_t)(t.rA * t.rx + t.rb), Rw = t.last ? 1 : (size_t)(t.rA2 * t.rx2 + t.rb2);
			// w_0 becomes c_0, kept to the end when it is also the last core
			W[j] = k == 0 && t.last ? DTensor({Lw, (size_t)t.n, Rw}) : DTensor::scratch({Lw, (size_t)t.n, Rw});
			t.w = W[j].data();
			asmT.push_back(t);
		}
		res_core_group(asmT, s);
		std::vector<GemmTask> g1, g2;
		for (size_t j = 0; j < J; ++j) {
			const Lock& L = P[j];
			if (k >= L.d) continue;
			const DTensor& Bc = L.b->cores[k];
			const size_t beta = Bc.dims[0], n = Bc.dims[1], beta2 = Bc.dims[2];
			double* gout;
			if (k + 1 == L.d) {
				gout = gfin.get() + j;
			} else {
				Gn[j] = DTensor({beta2, beta2});
				gout = Gn[j].data();
			}
			if (k == 0) {
				C[j] = std::move(W[j]);
				g1.push_back(gt(gemm_desc(1, 0, beta2, beta2, n), Bc.data(), Bc.data(), gout));
			} else {
				const size_t kk = Rk[j].dims[0], Lw = W[j].dims[0], Rw = W[j].dims[2];
				C[j] = k + 1 == L.d ? DTensor({kk, n, Rw}) : DTensor::scratch({kk, n, Rw});  // the last one is kept
				g1.push_back(gt(gemm_desc(0, 0, kk, n * Rw, Lw), Rk[j].data(), W[j].data(), C[j].data()));
				M[j] = DTensor::scratch({beta, n, beta2});
				g1.push_back(gt(gemm_desc(0, 0, beta, n * beta2, beta), G[j].data(), Bc.data(), M[j].data()));
				g2.push_back(gt(gemm_desc(1, 0, beta2, beta2, beta * n), Bc.data(), M[j].data(), gout));
			}
		}
		gemm_group_f64(g1, s);
		gemm_group_f64(g2, s);
		std::vector<RJob> qr;
		for (size_t j = 0; j < J; ++j) {
			const Lock& L = P[j];
			if (k >= L.d) continue;
			if (k + 1 == L.d) {
				last[j] = std::move(C[j]);
				continue;
			}
			const size_t cols = C[j].dims[2], rows = C[j].size() / cols;
			Rk[j] = DTensor({std::min(rows, cols), cols});
			qr.push_back(RJob{C[j].data(), (long long)rows, (long long)cols, Rk[j].data()});
			G[j] = std::move(Gn[j]);
		}
		hh_r_group(qr, s);
	}
	std::vector<Dot2Task> dt;
	for (size_t j = 0; j < J; ++j) dt.push_back(Dot2Task{(int)last[j].size(), last[j].data(), last[j].data(), nullptr, nullptr});
	DBuf<double> sums(2 * J);
	dot2_group(dt, sums.get(), s);
	std::vector<double> h(2 * J), gb(J);
	read_back(h.data(), sums.get(), 2 * J * sizeof(double), s);
	read_back(gb.data(), gfin.get(), J * sizeof(double), s);
	table_ring_release(s);
	for (size_t j = 0; j < J; ++j) out[j] = std::sqrt(std::max(0.0, h[2 * j])) / std::sqrt(gb[j]);
}

bool lockstep_supported(const ALSOptions& o, const std::vector<const TT*>& A) {
	if (o.sites != 1 || !o.assumeSPD || o.useResidualForEndCriterion || (o.precision != P_F64 && o.precision != P_F32)) return false;
	for (const TT* a : A) if (!a) return false;
	return true;
}

void als_round_lockstep(const ALSOptions& o, const std::vector<const TT*>& A, const std::vector<TT*>& x,
                        const std::vector<const TT*>& b, size_t sweeps, double eps, double* residuals, ALSStats* total) {
	const size_t count = x.size();
	std::vector<Lock> P(count);
	for (size_t j = 0; j < count; ++j) {
		XTT_REQUIRE(x[j]->order() > 0 && x[j]->dims == b[j]->dims && !x[j]->op && !b[j]->op, "lockstep: x and b must be TTTensors of equal dimensions");
		XTT_REQUIRE(A[j]->op && A[j]->order() == x[j]->order(), "lockstep: operator degree mismatch");
		require_correct_format(*x[j]);
		require_correct_format(*b[j]);
		require_correct_format(*A[j]);
		P[j].A = A[j]; P[j].x = x[j]; P[j].b = b[j]; P[j].d = x[j]->order();
	}
	hipStream_t s = stream();
	for (size_t sw = 0; sw < sweeps; ++sw) {
		als_lockstep(o, P, s);
		round_lockstep(P, eps, s);
	}
	if (residuals) residuals_lockstep(P, residuals, s);
	sync();
	if (total) {
		for (const Lock& L : P) {
			total->localSolves += L.localSolves;
			total->cgIterations += L.cgIters;
			total->refinements += L.refinements;
			total->halfSweeps += L.halfSweepsAll;
		}
	}
}

}  // namespace xtt
