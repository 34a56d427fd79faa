/*
 * Served path-tracing sessions (continuous batching) and the batched-launch
 * autotuner: host C++ over the library's own C ABI (include/lrhip.h).
 *
 * A session is a small path tracer bound to one resident scene: width x
 * height camera rays, then up to maxBounces bounce wavefronts, with one
 * shadow ray per path vertex traced as any-hit.  Each wavefront is shaded on
 * the GPU by lrh_shade (grey Lambertian surfaces lit by a square area light
 * through next-event estimation, cosine next vertices, Russian roulette)
 * or, in bounce mode, by lrh_rays_bounce.  Each session has one wavefront in
 * flight at a time.  This is the caller pattern of SLG's path tracer, which
 * enqueues one ray buffer per wavefront and consumes its hits
 * (src/slg/engines/pathocl/pathoclopenclthread.cpp:183-189); here the
 * wavefronts of many sessions share one launch.
 *
 * Every admitted session's buffers stay in HBM.  Each tick serves one cohort
 * (sessions are split over `cohorts` streams): it admits queued sessions
 * while the HBM budget and the per-tick ray budget allow, gathers the ready
 * wavefront of every live session of the cohort (closest-hit rays + the
 * shadow rays of the previous vertex) into ONE lrh_trace_batch launch, then
 * generates every session's next wavefront on the GPU (shade mode: one
 * lrh_shade_batch launch for the whole tick) and retires finished sessions.
 * Each cohort's launch takes 80% of the persistent grid unless the caller's
 * launch shape says otherwise, so the other cohort's tick runs beside it and
 * fills the CUs its waves leave while it drains.  Sessions that arrive while others are on the GPU join at their
 * cohort's next tick.  The host needs no result of a tick to build the next
 * one, so while one cohort's tick drains the other cohort's is already
 * queued; a tick completes when its end event does.
 *
 * Latency of a wavefront = completion of the tick that traced it minus the
 * moment it became ready (admission for a camera wavefront, the previous
 * tick's completion otherwise).  Busy time = the union of the ticks' GPU
 * intervals (events around each tick); trace time = the union of the
 * intervals of the batched traces alone.
 */
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <deque>
#include <fstream>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lrhip.h"
#include "lrh_internal.h"

namespace {

struct SrvError : std::runtime_error {
	int code;
	SrvError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void chk(int rc, const char *what) {
	if (rc != LRH_OK) throw SrvError(rc, std::string(what) + ": " + lrh_last_error());
}

template <class F>
int guard(F fn) {
	try {
		fn();
		return LRH_OK;
	} catch (const SrvError &e) {
		return lrhSetError(e.code, e.what());
	} catch (const std::bad_alloc &) {
		return lrhSetError(LRH_ENOMEM, "host allocation failed");
	} catch (const std::exception &e) {
		return lrhSetError(LRH_EINVAL, e.what());
	}
}

double clockSeconds() {
	return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// SplitMix64 (Steele, Lea, Flood 2014): the camera placement of a session
uint64_t splitmix(uint64_t &s) {
	uint64_t z = (s += 0x9e3779b97f4a7c15ull);
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
	return z ^ (z >> 31);
}
double uniform(uint64_t &s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

constexpr size_t kRay = LRH_RAY_BYTES, kHit = LRH_HIT_BYTES, kPath = sizeof(lrh_path);

struct Buffers {
	void *rays[2] = {nullptr, nullptr};
	void *shadow = nullptr, *hits = nullptr, *shadowHits = nullptr, *paths = nullptr;
};

struct Recorded {
	std::vector<uint8_t> rays[2], hits[2];  // [0] closest-hit wavefront, [1] shadow wavefront
	bool has[2] = {false, false};
};

struct Session {
	lrh_session_spec spec{};
	unsigned cohort = 0;
	void *stream = nullptr;
	Buffers b;
	unsigned cur = 0, step = 0;
	double ready = 0, admitted = 0;
	bool record = false;
	std::vector<Recorded> rec;
	std::vector<float> radiance;
};

struct Inflight {
	lrh_event *e0 = nullptr, *eMid = nullptr, *e1 = nullptr;
	std::vector<Session *> traced, finished;
};

struct Captured {
	unsigned scene;
	size_t offset;   // byte offset of the rays in the arena
	unsigned count, query;
};

double unionSeconds(std::vector<std::pair<double, double>> iv) {
	std::sort(iv.begin(), iv.end());
	double total = 0, end = -1e300;
	for (const auto &p : iv) {
		if (p.second <= end) continue;
		total += p.second - std::max(p.first, end);
		end = p.second;
	}
	return total / 1e3;
}

double percentile(std::vector<double> v, double q) {
	if (v.empty()) return 0.0;
	std::sort(v.begin(), v.end());
	// linear interpolation between closest ranks (numpy's default)
	const double pos = q / 100.0 * (double)(v.size() - 1);
	const size_t lo = (size_t)std::floor(pos), hi = std::min(lo + 1, v.size() - 1);
	return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

}  // namespace

struct lrh_server {
	lrh_device *dev = nullptr;
	std::vector<lrh_scene *> scenes;
	std::map<const lrh_scene *, unsigned long long> weight;  // records per scene: the cost order of a tick's items
	lrh_server_config cfg{};
	unsigned n = 0;                       // rays per wavefront (width * height)
	size_t sessionBytes = 0;
	std::vector<void *> streams;          // [0] = the device stream (NULL)
	std::vector<std::vector<Session *>> cohorts;
	std::vector<Inflight> inflight;
	std::vector<bool> busyCohort;
	unsigned turn = 0;
	std::deque<lrh_session_spec> queue;   // arrival: absolute clock seconds
	std::set<unsigned> recordSet, usedSids;
	std::vector<std::unique_ptr<Session>> sessions;
	std::map<unsigned, Session *> bySid;
	std::vector<Session *> done;
	std::vector<double> lat, qwait;
	std::vector<std::pair<double, double>> intervals, traceIntervals;
	std::vector<Buffers> pool;
	size_t allocated = 0;
	unsigned long long ticks = 0, wavefronts = 0;
	unsigned maxActive = 0;
	void *counter = nullptr;              // device u32: live rays written by the shading steps
	void *materials = nullptr, *lights = nullptr;
	unsigned nMaterials = 0, nLights = 0;
	lrh_event *ref = nullptr;
	std::vector<lrh_event *> evPool;
	std::vector<uint8_t> zeros;           // host zeros for clearing recorded hit buffers
	std::vector<lrh_batch_item> items;
	// capture of every traced wavefront (for the plain single-batch comparison)
	void *arena = nullptr;
	size_t arenaBytes = 0, arenaUsed = 0;
	bool captureOverflow = false;
	std::vector<Captured> captured;

	lrh_event *event() {
		if (!evPool.empty()) {
			lrh_event *e = evPool.back();
			evPool.pop_back();
			return e;
		}
		lrh_event *e = nullptr;
		chk(lrh_event_create(dev, &e), "lrh_event_create");
		return e;
	}
	void *alloc(size_t bytes) {
		void *p = nullptr;
		chk(lrh_alloc(dev, bytes, &p), "lrh_alloc");
		return p;
	}
	double now() const { return clockSeconds(); }
	size_t resident() const { return allocated - pool.size(); }

	Buffers takeBuffers() {
		if (!pool.empty()) {
			Buffers b = pool.back();
			pool.pop_back();
			return b;
		}
		Buffers b;
		b.rays[0] = alloc(n * kRay);
		b.rays[1] = alloc(n * kRay);
		b.shadow = alloc(n * kRay);
		b.hits = alloc(n * kHit);
		b.shadowHits = alloc(n * kHit);
		b.paths = alloc(n * kPath);
		++allocated;
		return b;
	}

	bool canAdmit(unsigned c) const {
		if ((resident() + 1) * sessionBytes > cfg.hbmBudget) return false;
		const unsigned long long rays = (unsigned long long)(cohorts[c].size() + 1) * 2ull * n;
		return rays <= std::max<unsigned long long>(cfg.maxRaysPerTick, 2ull * n);
	}

	void startSession(const lrh_session_spec &spec, unsigned c, double t) {
		std::unique_ptr<Session> s(new Session());
		s->spec = spec;
		s->cohort = c;
		s->stream = streams[c];
		s->b = takeBuffers();
		s->admitted = s->ready = t;
		s->record = recordSet.count(spec.sid) != 0;
		// camera: eye in the upper part of the scene box, looking at its lower middle
		uint64_t rs = spec.seed ^ 0x5851f42d4c957f2dull;
		const double ext = cfg.extent;
		lrh_camera cam{};
		for (int k = 0; k < 3; ++k) cam.eye[k] = (float)(ext * (0.1 + 0.8 * uniform(rs)));
		cam.eye[2] = (float)(ext * (0.6 + 0.35 * uniform(rs)));
		cam.look[0] = (float)(ext * 0.5 + ext * (0.2 * uniform(rs) - 0.1));
		cam.look[1] = (float)(ext * 0.5 + ext * (0.2 * uniform(rs) - 0.1));
		cam.look[2] = (float)(ext * 0.3 + ext * (0.2 * uniform(rs) - 0.1));
		cam.up[2] = 1.f;
		cam.vfovDeg = 50.f;
		cam.mint = 1e-4f;
		cam.width = cfg.width;
		cam.height = cfg.height;
		chk(lrh_rays_camera(dev, &cam, s->b.rays[0], s->stream), "lrh_rays_camera");
		if (cfg.mode == LRH_SERVE_SHADE) chk(lrh_paths_init(dev, s->b.paths, n, s->stream), "lrh_paths_init");
		qwait.push_back(t - spec.arrival);
		cohorts[c].push_back(s.get());
		bySid[spec.sid] = s.get();
		sessions.push_back(std::move(s));
	}

	void admit(double t) {
		while (!queue.empty() && queue.front().arrival <= t) {
			unsigned c = 0;
			for (unsigned k = 1; k < cohorts.size(); ++k)
				if (cohorts[k].size() < cohorts[c].size()) c = k;
			if (!canAdmit(c)) return;
			const lrh_session_spec spec = queue.front();
			queue.pop_front();
			startSession(spec, c, t);
		}
	}

	// batch items of a session's current wavefront
	void sessionItems(Session *s) {
		lrh_scene *sc = scenes[s->spec.scene];
		if (s->step <= cfg.maxBounces) items.push_back({sc, s->b.rays[s->cur], s->b.hits, n, LRH_QUERY_CLOSEST});
		if (s->step >= 1) items.push_back({sc, s->b.shadow, s->b.shadowHits, n, LRH_QUERY_ANY});
	}

	// after the tick's trace: the session's next wavefront, generated on the
	// GPU (shade mode: an item of the tick's one lrh_shade_batch launch)
	bool advance(Session *s, std::vector<lrh_shade_item> &shade) {
		lrh_scene *sc = scenes[s->spec.scene];
		const unsigned last = cfg.maxBounces;
		if (cfg.mode == LRH_SERVE_SHADE) {
			// vertex `step` (its shadow rays' hits come from the previous
			// vertex); one more step after the last vertex only gathers the
			// light its shadow rays found
			const unsigned nxt = 1 - s->cur;
			lrh_shade_item it{};
			it.scene = sc;
			it.rays = s->b.rays[s->cur];
			it.hits = s->b.hits;
			it.shadowHits = s->step >= 1 ? s->b.shadowHits : nullptr;
			it.paths = s->b.paths;
			it.nextRays = s->b.rays[nxt];
			it.shadowRays = s->b.shadow;
			it.params.count = n;
			it.params.maxDepth = last + 1;
			it.params.rrDepth = cfg.rrStart + 1;
			it.params.rrSurvive = cfg.rrSurvive;
			it.params.seed = s->spec.seed;
			it.params.pixelBase = 0;
			shade.push_back(it);
			s->cur = nxt;
		} else if (s->step <= last) {
			const unsigned nxt = 1 - s->cur;
			chk(lrh_rays_bounce(sc, s->b.rays[s->cur], s->b.hits, n, s->step, s->spec.seed, 0, cfg.rrStart,
					cfg.rrSurvive, cfg.light, s->b.rays[nxt], s->b.shadow, (unsigned *)counter, s->stream),
					"lrh_rays_bounce");
			s->cur = nxt;
		}
		++s->step;
		return s->step <= last + 1;
	}

	void retire(unsigned c, bool block) {
		if (!busyCohort[c]) return;
		Inflight &f = inflight[c];
		if (!block) {
			const int q = lrh_event_query(f.e1);
			if (q < 0) chk(q, "lrh_event_query");
			if (q == 0) return;
		}
		chk(lrh_event_sync(f.e1), "lrh_event_sync");
		const double t = now();
		float a = 0, m = 0, b = 0;
		chk(lrh_event_elapsed(ref, f.e0, &a), "lrh_event_elapsed");
		chk(lrh_event_elapsed(ref, f.eMid, &m), "lrh_event_elapsed");
		chk(lrh_event_elapsed(ref, f.e1, &b), "lrh_event_elapsed");
		intervals.push_back({a, b});
		traceIntervals.push_back({a, m});
		for (Session *s : f.traced) {
			lat.push_back(t - s->ready);
			s->ready = t;
		}
		for (Session *s : f.finished) {
			if (s->record && cfg.mode == LRH_SERVE_SHADE) {
				std::vector<lrh_path> paths(n);
				chk(lrh_memcpy(dev, paths.data(), s->b.paths, n * kPath, LRH_D2H, s->stream), "lrh_memcpy");
				s->radiance.resize(3 * (size_t)n);
				for (unsigned i = 0; i < n; ++i)
					for (int k = 0; k < 3; ++k) s->radiance[3 * (size_t)i + k] = paths[i].rad[k];
			}
			pool.push_back(s->b);
			s->b = Buffers();
			done.push_back(s);
		}
		evPool.push_back(f.e0);
		evPool.push_back(f.eMid);
		evPool.push_back(f.e1);
		f = Inflight();
		busyCohort[c] = false;
	}

	void drain() {
		for (unsigned c = 0; c < cohorts.size(); ++c) retire(c, true);
	}

	bool tick() {
		const unsigned c = turn;
		turn = (c + 1) % (unsigned)cohorts.size();
		retire(c, true);
		for (unsigned k = 0; k < cohorts.size(); ++k) retire(k, false);
		admit(now());
		std::vector<Session *> &coh = cohorts[c];
		if (coh.empty()) return false;
		unsigned active = 0;
		for (const auto &x : cohorts) active += (unsigned)x.size();
		maxActive = std::max(maxActive, active);
		void *st = streams[c];
		items.clear();
		for (Session *s : coh) sessionItems(s);
		// the costliest rays take the first tickets (closest-hit rays before
		// shadow rays, larger scenes first), so the end of the launch is
		// spent on short rays and its tail is short
		std::stable_sort(items.begin(), items.end(), [this](const lrh_batch_item &a, const lrh_batch_item &b) {
			if (a.query != b.query) return a.query < b.query;
			return weight.at(a.scene) > weight.at(b.scene);
		});
		bool anyRecord = false;
		for (Session *s : coh) {
			if (!s->record) continue;
			// masked rays leave their hit records untouched: clear the recorded
			// sessions' hit buffers so what is compared does not depend on the
			// buffers' earlier contents
			chk(lrh_memcpy_async(dev, s->b.hits, zeros.data(), n * kHit, LRH_H2D, st), "lrh_memcpy_async");
			chk(lrh_memcpy_async(dev, s->b.shadowHits, zeros.data(), n * kHit, LRH_H2D, st), "lrh_memcpy_async");
			anyRecord = true;
		}
		Inflight &f = inflight[c];
		f.e0 = event();
		f.eMid = event();
		f.e1 = event();
		chk(lrh_event_record(f.e0, st), "lrh_event_record");
		chk(lrh_trace_batch(dev, items.data(), (unsigned)items.size(), &cfg.opts, st), "lrh_trace_batch");
		chk(lrh_event_record(f.eMid, st), "lrh_event_record");
		wavefronts += coh.size();
		if (anyRecord) {
			chk(lrh_sync(dev, st), "lrh_sync");
			for (Session *s : coh) {
				if (!s->record) continue;
				Recorded r;
				for (int k = 0; k < 2; ++k) {
					const bool traced = k == 0 ? s->step <= cfg.maxBounces : s->step >= 1;
					if (!traced) continue;
					r.has[k] = true;
					r.rays[k].resize(n * kRay);
					r.hits[k].resize(n * kHit);
					chk(lrh_memcpy(dev, r.rays[k].data(), k == 0 ? s->b.rays[s->cur] : s->b.shadow, n * kRay,
							LRH_D2H, st), "lrh_memcpy");
					chk(lrh_memcpy(dev, r.hits[k].data(), k == 0 ? s->b.hits : s->b.shadowHits, n * kHit,
							LRH_D2H, st), "lrh_memcpy");
				}
				s->rec.push_back(std::move(r));
			}
		}
		if (arena) {
			// copy of every ray buffer this tick traced, before the shading
			// step overwrites the shadow rays
			for (const lrh_batch_item &it : items) {
				const size_t bytes = (size_t)it.count * kRay;
				if (arenaUsed + bytes > arenaBytes) {
					captureOverflow = true;
					continue;
				}
				chk(lrh_memcpy_async(dev, (uint8_t *)arena + arenaUsed, it.rays, bytes, LRH_D2D, st),
						"lrh_memcpy_async");
				unsigned si = 0;
				while (scenes[si] != it.scene) ++si;
				captured.push_back({si, arenaUsed, it.count, it.query});
				arenaUsed += (bytes + 255) & ~(size_t)255;
			}
		}
		std::vector<Session *> still, finished;
		std::vector<lrh_shade_item> shade;
		for (Session *s : coh) (advance(s, shade) ? still : finished).push_back(s);
		for (size_t k = 0; k < shade.size(); k += LRH_SHADE_BATCH_MAX)
			chk(lrh_shade_batch(shade.data() + k, (unsigned)std::min<size_t>(LRH_SHADE_BATCH_MAX, shade.size() - k),
					(const lrh_material *)materials, nMaterials, (const lrh_light_tri *)lights, nLights,
					(unsigned *)counter, st), "lrh_shade_batch");
		chk(lrh_event_record(f.e1, st), "lrh_event_record");
		f.traced = coh;
		f.finished = finished;
		busyCohort[c] = true;
		coh = still;
		++ticks;
		return true;
	}

	bool pending() const {
		if (!queue.empty()) return true;
		for (const auto &c : cohorts)
			if (!c.empty()) return true;
		for (bool b : busyCohort)
			if (b) return true;
		return false;
	}

	void release() {
		for (Buffers &b : pool)
			for (void *p : {b.rays[0], b.rays[1], b.shadow, b.hits, b.shadowHits, b.paths}) (void)lrh_free(dev, p);
		pool.clear();
		for (auto &s : sessions)
			if (s->b.paths)
				for (void *p : {s->b.rays[0], s->b.rays[1], s->b.shadow, s->b.hits, s->b.shadowHits, s->b.paths})
					(void)lrh_free(dev, p);
		for (size_t k = 1; k < streams.size(); ++k) (void)lrh_stream_destroy(dev, streams[k]);
		streams.resize(1);
		for (lrh_event *e : evPool) (void)lrh_event_destroy(e);
		evPool.clear();
		for (Inflight &f : inflight)
			for (lrh_event *e : {f.e0, f.eMid, f.e1})
				if (e) (void)lrh_event_destroy(e);
		inflight.clear();
		if (ref) (void)lrh_event_destroy(ref);
		ref = nullptr;
		for (void *p : {counter, materials, lights, arena})
			if (p) (void)lrh_free(dev, p);
		counter = materials = lights = arena = nullptr;
	}
};

extern "C" {

int lrh_server_create(lrh_device *dev, lrh_scene *const *scenes, unsigned int nScenes, const lrh_server_config *cfg,
		lrh_server **out) {
	if (!dev || !out || !scenes || nScenes == 0) return lrhSetError(LRH_EINVAL, "null argument or no scenes");
	*out = nullptr;
	std::unique_ptr<lrh_server> s(new lrh_server());
	const int rc = guard([&]() {
		s->dev = dev;
		for (unsigned i = 0; i < nScenes; ++i) {
			if (!scenes[i]) throw SrvError(LRH_EINVAL, "null scene " + std::to_string(i));
			s->scenes.push_back(scenes[i]);
			lrh_scene_info info{};
			chk(lrh_scene_get_info(scenes[i], &info), "lrh_scene_get_info");
			s->weight[scenes[i]] = info.primCount;
		}
		lrh_server_config c{};
		if (cfg) c = *cfg;
		if (!c.width) c.width = 512;
		if (!c.height) c.height = 512;
		if (!c.maxBounces) c.maxBounces = 8;
		if (!c.rrStart) c.rrStart = 2;
		if (!(c.rrSurvive > 0.f)) c.rrSurvive = 0.7f;
		if (!c.hbmBudget) c.hbmBudget = 4ull << 30;
		if (!c.maxRaysPerTick) c.maxRaysPerTick = 3u << 20;
		if (!c.cohorts) c.cohorts = 2;
		// each cohort's launch takes 80% of the persistent grid, so the other
		// cohort's tick starts on the rest and takes over the CUs its waves
		// leave while it drains (a tick carries about one live ray per lane
		// of a full grid: its drain is a large part of it)
		if (!c.opts.gridPercent && !c.opts.blocksPerCU && c.cohorts > 1) c.opts.gridPercent = 80;
		if (!(c.extent > 0.f)) c.extent = 100.f;
		if (c.light[0] == 0.f && c.light[1] == 0.f && c.light[2] == 0.f) {
			c.light[0] = 0.5f * c.extent;
			c.light[1] = 0.5f * c.extent;
			c.light[2] = 1.5f * c.extent;
		}
		if (c.mode > LRH_SERVE_BOUNCE) throw SrvError(LRH_EINVAL, "mode must be LRH_SERVE_SHADE or LRH_SERVE_BOUNCE");
		if (c.cohorts > 8) throw SrvError(LRH_EINVAL, "at most 8 cohorts");
		if ((unsigned long long)c.width * c.height > (1ull << 26)) throw SrvError(LRH_ERANGE, "session image too large");
		s->cfg = c;
		s->n = c.width * c.height;
		s->sessionBytes = 3 * (size_t)s->n * kRay + 2 * (size_t)s->n * kHit + (size_t)s->n * kPath;
		s->streams.push_back(nullptr);
		for (unsigned k = 1; k < c.cohorts; ++k) {
			void *st = nullptr;
			chk(lrh_stream_create(dev, &st), "lrh_stream_create");
			s->streams.push_back(st);
		}
		s->cohorts.resize(c.cohorts);
		s->inflight.resize(c.cohorts);
		s->busyCohort.assign(c.cohorts, false);
		s->counter = s->alloc(8);
		const uint32_t z2[2] = {0, 0};
		chk(lrh_memcpy(dev, s->counter, z2, 8, LRH_H2D, nullptr), "lrh_memcpy");
		s->zeros.assign((size_t)s->n * kHit, 0);
		if (c.mode == LRH_SERVE_SHADE) {
			// grey Lambertian surfaces (one entry per DataSet mesh of any scene)
			// and a square light facing down from the light centre, 0.3 x extent
			// wide, radiance 30
			std::vector<lrh_material> m(1024);
			for (auto &x : m) {
				x.kd[0] = x.kd[1] = x.kd[2] = 0.6f;
				x.le[0] = x.le[1] = x.le[2] = 0.f;
			}
			s->nMaterials = (unsigned)m.size();
			s->materials = s->alloc(m.size() * sizeof(lrh_material));
			chk(lrh_memcpy(dev, s->materials, m.data(), m.size() * sizeof(lrh_material), LRH_H2D, nullptr), "lrh_memcpy");
			const float cx = c.light[0], cy = c.light[1], h = c.light[2], w = 0.15f * c.extent;
			const float q[4][3] = {{cx - w, cy - w, h}, {cx + w, cy - w, h}, {cx + w, cy + w, h}, {cx - w, cy + w, h}};
			const int tri[2][3] = {{0, 2, 1}, {0, 3, 2}};
			lrh_light_tri lt[2];
			for (int t = 0; t < 2; ++t)
				for (int k = 0; k < 3; ++k) {
					lt[t].p0[k] = q[tri[t][0]][k];
					lt[t].e1[k] = q[tri[t][1]][k] - q[tri[t][0]][k];
					lt[t].e2[k] = q[tri[t][2]][k] - q[tri[t][0]][k];
					lt[t].le[k] = 30.f;
				}
			s->nLights = 2;
			s->lights = s->alloc(sizeof(lt));
			chk(lrh_memcpy(dev, s->lights, lt, sizeof(lt), LRH_H2D, nullptr), "lrh_memcpy");
		}
		chk(lrh_sync(dev, nullptr), "lrh_sync");
		chk(lrh_event_create(dev, &s->ref), "lrh_event_create");
		chk(lrh_event_record(s->ref, nullptr), "lrh_event_record");
	});
	if (rc != LRH_OK) {
		s->release();
		return rc;
	}
	*out = s.release();
	return LRH_OK;
}

int lrh_server_submit(lrh_server *s, const lrh_session_spec *specs, unsigned int n) {
	if (!s || (n && !specs)) return lrhSetError(LRH_EINVAL, "null argument");
	return guard([&]() {
		const double t = s->now();
		std::vector<lrh_session_spec> v(specs, specs + n);
		std::set<unsigned> ids;
		for (auto &x : v) {
			if (x.scene >= s->scenes.size()) throw SrvError(LRH_EINVAL, "session scene index out of range");
			if (s->usedSids.count(x.sid) || !ids.insert(x.sid).second)
				throw SrvError(LRH_EINVAL, "session id already used");
			x.arrival = t + std::max(0.0, x.arrival);
		}
		s->usedSids.insert(ids.begin(), ids.end());
		for (auto &x : v) s->queue.push_back(x);
		std::stable_sort(s->queue.begin(), s->queue.end(),
				[](const lrh_session_spec &a, const lrh_session_spec &b) { return a.arrival < b.arrival; });
	});
}

int lrh_server_record(lrh_server *s, unsigned int sid) {
	if (!s) return lrhSetError(LRH_EINVAL, "null server");
	s->recordSet.insert(sid);
	return LRH_OK;
}

int lrh_server_tick(lrh_server *s, int *traced) {
	if (!s) return lrhSetError(LRH_EINVAL, "null server");
	return guard([&]() {
		const bool t = s->tick();
		if (traced) *traced = t ? 1 : 0;
	});
}

int lrh_server_run(lrh_server *s, double timeLimit, double *wallSeconds) {
	if (!s) return lrhSetError(LRH_EINVAL, "null server");
	return guard([&]() {
		const double t0 = s->now();
		unsigned idle = 0;
		while (s->pending() && s->now() - t0 < timeLimit) {
			if (s->tick()) {
				idle = 0;
				continue;
			}
			if (++idle < s->cohorts.size()) continue;
			idle = 0;
			bool busy = false;
			for (bool b : s->busyCohort) busy = busy || b;
			if (busy) {
				s->drain();
			} else if (!s->queue.empty()) {
				// nothing admitted yet: wait for the next arrival
				const double dt = s->queue.front().arrival - s->now();
				if (dt > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(dt, 1e-3)));
			}
		}
		s->drain();
		if (wallSeconds) *wallSeconds = s->now() - t0;
	});
}

int lrh_server_get_stats(lrh_server *s, lrh_server_stats *st) {
	if (!s || !st) return lrhSetError(LRH_EINVAL, "null argument");
	return guard([&]() {
		s->drain();
		std::memset(st, 0, sizeof(*st));
		uint32_t c[2] = {0, 0};
		chk(lrh_memcpy(s->dev, c, s->counter, 8, LRH_D2H, nullptr), "lrh_memcpy");
		st->sessionsDone = s->done.size();
		st->sessionsQueued = s->queue.size();
		st->ticks = s->ticks;
		st->wavefronts = s->wavefronts;
		st->liveRays = (unsigned long long)c[0] + (unsigned long long)s->done.size() * s->n;
		st->maxActive = s->maxActive;
		st->sessionBuffersAllocated = (unsigned)s->allocated;
		st->busySeconds = unionSeconds(s->intervals);
		st->traceSeconds = unionSeconds(s->traceIntervals);
		st->latencySamples = s->lat.size();
		std::vector<double> ms(s->lat.size());
		for (size_t i = 0; i < ms.size(); ++i) ms[i] = s->lat[i] * 1e3;
		st->p50Ms = percentile(ms, 50);
		st->p99Ms = percentile(ms, 99);
		st->maxMs = ms.empty() ? 0.0 : *std::max_element(ms.begin(), ms.end());
		std::vector<double> qw(s->qwait.size());
		for (size_t i = 0; i < qw.size(); ++i) qw[i] = s->qwait[i] * 1e3;
		st->queueP99Ms = percentile(qw, 99);
	});
}

int lrh_server_latencies(lrh_server *s, double *ms, size_t cap, size_t *n) {
	if (!s || (cap && !ms)) return lrhSetError(LRH_EINVAL, "null argument");
	const size_t k = std::min(cap, s->lat.size());
	for (size_t i = 0; i < k; ++i) ms[i] = s->lat[i] * 1e3;
	if (n) *n = s->lat.size();
	return LRH_OK;
}

int lrh_server_recorded(lrh_server *s, unsigned int sid, unsigned int step, unsigned int kind, void *rays, void *hits,
		unsigned int *count) {
	if (!s || kind > 1) return lrhSetError(LRH_EINVAL, "null server or bad kind");
	auto it = s->bySid.find(sid);
	if (it == s->bySid.end() || !it->second->record) return lrhSetError(LRH_EINVAL, "session was not recorded");
	const Session *x = it->second;
	if (step >= x->rec.size() || !x->rec[step].has[kind])
		return lrhSetError(LRH_EINVAL, "no wavefront of that kind at that step");
	if (rays) std::memcpy(rays, x->rec[step].rays[kind].data(), x->rec[step].rays[kind].size());
	if (hits) std::memcpy(hits, x->rec[step].hits[kind].data(), x->rec[step].hits[kind].size());
	if (count) *count = s->n;
	return LRH_OK;
}

int lrh_server_radiance(lrh_server *s, unsigned int sid, float *rgb, unsigned int count) {
	if (!s || !rgb) return lrhSetError(LRH_EINVAL, "null argument");
	auto it = s->bySid.find(sid);
	if (it == s->bySid.end() || it->second->radiance.empty())
		return lrhSetError(LRH_EINVAL, "no radiance recorded for that session");
	if (count != s->n) return lrhSetError(LRH_EINVAL, "count must be width * height");
	std::memcpy(rgb, it->second->radiance.data(), it->second->radiance.size() * sizeof(float));
	return LRH_OK;
}

int lrh_server_capture(lrh_server *s, unsigned long long maxBytes) {
	if (!s || maxBytes == 0) return lrhSetError(LRH_EINVAL, "null server or zero bytes");
	if (s->arena) return lrhSetError(LRH_EINVAL, "capture already enabled");
	if (!s->sessions.empty()) return lrhSetError(LRH_EINVAL, "enable capture before the first session is admitted");
	return guard([&]() {
		s->arena = s->alloc(maxBytes);
		s->arenaBytes = maxBytes;
	});
}

int lrh_server_plain_batch(lrh_server *s, const lrh_trace_opts *opts, unsigned int reps, lrh_plain_batch_result *res) {
	if (!s || !res) return lrhSetError(LRH_EINVAL, "null argument");
	if (!s->arena) return lrhSetError(LRH_EINVAL, "nothing captured (lrh_server_capture)");
	if (s->captureOverflow) return lrhSetError(LRH_ERANGE, "capture arena overflowed: give lrh_server_capture more bytes");
	if (s->captured.empty()) return lrhSetError(LRH_EINVAL, "no wavefront was captured");
	if (s->captured.size() > 4096) return lrhSetError(LRH_ERANGE, "more than 4096 captured buffers");
	return guard([&]() {
		s->drain();
		std::memset(res, 0, sizeof(*res));
		unsigned long long total = 0;
		for (const Captured &c : s->captured) total += c.count;
		void *hits = s->alloc(total * kHit);
		std::vector<lrh_batch_item> items;
		unsigned long long off = 0;
		for (const Captured &c : s->captured) {
			items.push_back({s->scenes[c.scene], (uint8_t *)s->arena + c.offset, (uint8_t *)hits + off * kHit, c.count,
					c.query});
			off += c.count;
		}
		const unsigned r = std::max(1u, reps);
		lrh_event *a = nullptr, *b = nullptr;
		chk(lrh_event_create(s->dev, &a), "lrh_event_create");
		chk(lrh_event_create(s->dev, &b), "lrh_event_create");
		auto timeIt = [&](auto fn) {
			fn();  // warm-up
			chk(lrh_event_record(a, nullptr), "lrh_event_record");
			for (unsigned k = 0; k < r; ++k) fn();
			chk(lrh_event_record(b, nullptr), "lrh_event_record");
			chk(lrh_event_sync(b), "lrh_event_sync");
			float ms = 0;
			chk(lrh_event_elapsed(a, b, &ms), "lrh_event_elapsed");
			return (double)ms / r;
		};
		// default: the served shape over the whole grid (nothing runs beside a plain batch)
		lrh_trace_opts po = s->cfg.opts;
		po.gridPercent = 0;
		const lrh_trace_opts *o = opts ? opts : &po;
		res->batchMs = timeIt([&]() {
			chk(lrh_trace_batch(s->dev, items.data(), (unsigned)items.size(), o, nullptr), "lrh_trace_batch");
		});
		res->perBufferMs = timeIt([&]() {
			for (const lrh_batch_item &it : items)
				chk(lrh_trace_ex(it.scene, it.rays, it.hits, it.count, it.query, o, nullptr), "lrh_trace_ex");
		});
		(void)lrh_event_destroy(a);
		(void)lrh_event_destroy(b);
		(void)lrh_free(s->dev, hits);
		uint32_t c[2] = {0, 0};
		chk(lrh_memcpy(s->dev, c, s->counter, 8, LRH_D2H, nullptr), "lrh_memcpy");
		res->buffers = (unsigned)items.size();
		res->rays = total;
		res->liveRays = (unsigned long long)c[0] + (unsigned long long)s->done.size() * s->n;
	});
}

int lrh_server_destroy(lrh_server *s) {
	if (!s) return LRH_OK;
	const int rc = guard([&]() { s->drain(); });
	(void)lrh_sync(s->dev, nullptr);
	for (size_t k = 1; k < s->streams.size(); ++k) (void)lrh_sync(s->dev, s->streams[k]);
	s->release();
	delete s;
	return rc;
}

// ---------------------------------------------------------------------------
// autotuner of the batched launch shape, with an on-disk cache
// ---------------------------------------------------------------------------
int lrh_autotune_batch(lrh_device *dev, const lrh_batch_item *items, unsigned int nItems,
		const lrh_trace_opts *candidates, unsigned int nCandidates, const char *cachePath, const char *key,
		lrh_trace_opts *best, float *bestMs, float *candidateMs) {
	if (!dev || !items || !nItems || !best) return lrhSetError(LRH_EINVAL, "null argument");
	if (candidates && !nCandidates) return lrhSetError(LRH_EINVAL, "empty candidate list");
	return guard([&]() {
		const std::string k = key ? key : "";
		if (k.find_first_of(" \t\n") != std::string::npos) throw SrvError(LRH_EINVAL, "key must not contain spaces");
		std::vector<lrh_trace_opts> cand;
		if (candidates) {
			cand.assign(candidates, candidates + nCandidates);
		} else {
			for (unsigned blk : {128u, 256u})
				for (unsigned lds : {4u, 8u, 16u})
					for (unsigned ref : {4u, 8u, 12u, 16u, 24u}) {
						lrh_trace_opts o{};
						o.blockSize = blk;
						o.ldsStack = lds;
						o.refillThreshold = ref;
						cand.push_back(o);
					}
		}
		// cache file: one line per key, "key block lds refill ms blocksPerCU
		// leafTris gridPercent" (the last three may be absent: zero)
		if (cachePath && !k.empty()) {
			std::ifstream in(cachePath);
			std::string line;
			while (std::getline(in, line)) {
				std::istringstream ls(line);
				std::string kk;
				unsigned blk = 0, lds = 0, ref = 0;
				float ms = 0;
				if (ls >> kk >> blk >> lds >> ref >> ms && kk == k) {
					unsigned bpc = 0, leaf = 0, pct = 0;
					if (ls >> bpc && ls >> leaf) ls >> pct;
					std::memset(best, 0, sizeof(*best));
					best->blockSize = blk;
					best->ldsStack = lds;
					best->refillThreshold = ref;
					best->blocksPerCU = bpc;
					best->leafTris = leaf;
					best->gridPercent = pct;
					if (bestMs) *bestMs = ms;
					if (candidateMs)
						for (size_t i = 0; i < cand.size(); ++i) candidateMs[i] = NAN;
					return;
				}
			}
		}
		lrh_event *a = nullptr, *b = nullptr;
		chk(lrh_event_create(dev, &a), "lrh_event_create");
		chk(lrh_event_create(dev, &b), "lrh_event_create");
		float bestT = INFINITY;
		lrh_trace_opts bo{};
		for (size_t ci = 0; ci < cand.size(); ++ci) {
			const lrh_trace_opts &o = cand[ci];
			if (candidateMs) candidateMs[ci] = NAN;
			lrh_batch *bt = nullptr;
			// a shape the device refuses (LDS limit) is skipped
			if (lrh_batch_create(dev, items, nItems, &o, &bt) != LRH_OK) continue;
			std::vector<float> t;
			bool ok = true;
			for (int rep = 0; rep < 5 && ok; ++rep) {
				ok = lrh_event_record(a, nullptr) == LRH_OK && lrh_batch_launch(bt, nullptr) == LRH_OK &&
						lrh_event_record(b, nullptr) == LRH_OK && lrh_event_sync(b) == LRH_OK;
				float ms = 0;
				if (ok && rep >= 2 && lrh_event_elapsed(a, b, &ms) == LRH_OK) t.push_back(ms);
			}
			(void)lrh_batch_destroy(bt);
			if (!ok) {
				const std::string why = lrh_last_error();
				(void)lrh_event_destroy(a);
				(void)lrh_event_destroy(b);
				throw SrvError(LRH_EHIP, "autotune launch failed: " + why);
			}
			std::sort(t.begin(), t.end());
			const float med = t[t.size() / 2];
			if (candidateMs) candidateMs[ci] = med;
			if (med < bestT) {
				bestT = med;
				bo = o;
			}
		}
		(void)lrh_event_destroy(a);
		(void)lrh_event_destroy(b);
		if (!(bestT < INFINITY)) throw SrvError(LRH_EINVAL, "no launch shape could be created");
		*best = bo;
		if (bestMs) *bestMs = bestT;
		if (cachePath && !k.empty()) {
			std::vector<std::string> keep;
			{
				std::ifstream in(cachePath);
				std::string line;
				while (std::getline(in, line)) {
					std::istringstream ls(line);
					std::string kk;
					if (ls >> kk && kk != k) keep.push_back(line);
				}
			}
			const std::string tmp = std::string(cachePath) + ".tmp";
			{
				std::ofstream out(tmp);
				if (!out) throw SrvError(LRH_EINVAL, std::string("cannot write the autotune cache ") + tmp);
				for (const auto &l : keep) out << l << "\n";
				out << k << " " << bo.blockSize << " " << bo.ldsStack << " " << bo.refillThreshold << " " << bestT << " "
					<< bo.blocksPerCU << " " << bo.leafTris << " " << bo.gridPercent << "\n";
			}
			if (std::rename(tmp.c_str(), cachePath) != 0)
				throw SrvError(LRH_EINVAL, std::string("cannot replace the autotune cache ") + cachePath);
		}
	});
}

int lrh_autotune_build(lrh_device *dev, const lrh_mesh *const *sceneMeshes, const unsigned int *sceneMeshCount,
		unsigned int nScenes, const lrh_batch_item *items, const unsigned int *itemScene, unsigned int nItems,
		const lrh_build_opts *candidates, unsigned int nCandidates, const lrh_trace_opts *shapes,
		unsigned int nShapes, unsigned int *bestCandidate, lrh_trace_opts *bestShape, float *candidateMs) {
	if (!dev || !sceneMeshes || !sceneMeshCount || !nScenes || !items || !itemScene || !nItems || !candidates ||
			!nCandidates || !bestCandidate || !bestShape)
		return lrhSetError(LRH_EINVAL, "null argument");
	if (shapes && !nShapes) return lrhSetError(LRH_EINVAL, "empty shape list");
	for (unsigned i = 0; i < nItems; ++i)
		if (itemScene[i] >= nScenes) return lrhSetError(LRH_EINVAL, "item " + std::to_string(i) + ": scene index out of range");
	return guard([&]() {
		float bestT = INFINITY;
		unsigned bestC = 0;
		lrh_trace_opts bestO{};
		std::vector<lrh_scene *> scs(nScenes, nullptr);
		auto release = [&]() {
			for (auto &sc : scs)
				if (sc) {
					(void)lrh_scene_destroy(sc);
					sc = nullptr;
				}
		};
		for (unsigned c = 0; c < nCandidates; ++c) {
			if (candidateMs) candidateMs[c] = NAN;
			for (unsigned j = 0; j < nScenes; ++j) {
				if (lrh_scene_create_bvh(dev, sceneMeshes[j], sceneMeshCount[j], &candidates[c], &scs[j]) != LRH_OK) {
					const std::string why = lrh_last_error();
					release();
					throw SrvError(LRH_EINVAL, "candidate " + std::to_string(c) + ", scene " + std::to_string(j) + ": " + why);
				}
			}
			std::vector<lrh_batch_item> it(items, items + nItems);
			for (unsigned i = 0; i < nItems; ++i) it[i].scene = scs[itemScene[i]];
			lrh_trace_opts o{};
			float ms = INFINITY;
			const int rc = lrh_autotune_batch(dev, it.data(), nItems, shapes, shapes ? nShapes : 0, nullptr, nullptr,
					&o, &ms, nullptr);
			release();
			if (rc != LRH_OK) throw SrvError(rc, std::string("candidate ") + std::to_string(c) + ": " + lrh_last_error());
			if (candidateMs) candidateMs[c] = ms;
			if (ms < bestT) {
				bestT = ms;
				bestC = c;
				bestO = o;
			}
		}
		*bestCandidate = bestC;
		*bestShape = bestO;
	});
}

}  // extern "C"
