/* SPDX-License-Identifier: MIT
 *
 * TEST INFRASTRUCTURE: part of the oracle (included at the end of oracle_modes.c), never linked into the product.
 *
 * The master's side of a ballpark restated in scalar C, one world at a time in the reference's object structure
 * (src/exact_ballpark/space.py, SpaceBallpark with is_master; src/exact_ballpark/full.py, FullBallpark on the
 * server): the top boxes that exist and their bubbles (Ballpark::InsertInBoxes with NearbyBubbles, space.py:729-817,
 * 896-906), the bubble dictionaries in insertion order and the interactive and keep-alive sets (AddToBubble,
 * space.py:922-966), InitializeBubbles with its seed fill (space.py:986-1015), UpdateBallBubble (:1017-1033),
 * CopyBubbles (:1035-1051), the dead bubbles (InDeadBubble, :974-984), RemoveBall's partition side (:1286-1317),
 * the proximity sensors, notification ranges and target tracking in libstdc++'s unordered_map order
 * (HandleProximities and its checks, :1345-1522; StdHashMap, :193-320), glibc rand() for shuffled sensors
 * (space.py's GlibcRandom), and the queries CheckVisibility, GetBallIdsInRange, GetBubbleAtCoordinates and
 * GetRemoteFollowers (:1057-1072, 1524-1610).
 *
 * The boxes below the tops are not kept: a box exists while a ball is filed in it or below it, so the walk of
 * GetNearbyBalls follows from the filings (oracle_modes.c's box set). The top boxes' table holds every top some
 * filing reaches, each with its bubble (-1 for none).
 *
 * What pins it: tests/test_oracle_space_golden.py holds every structure this file produces to the vectors the
 * reference's own FullBallpark(is_master=True) wrote (tests/golden/make_golden_space.py).
 */

#define MAX_TOPS 1024
#define MAX_ENTRIES 1024
#define MAX_DICTS 256
#define MAX_PAIRS 1024
#define MAX_MEVENTS 4096
#define MAX_GRAVE 256

enum { EV_PROXIMITY = 1, EV_RANGE = 2, EV_TRACKING = 3, EV_FINAL_CLEANUP = 4 };

typedef struct { int n; int64_t key[MAX_BALLS]; uint64_t bkt, next; } OUMapBig;   /* the proximity balls */

typedef struct OMaster {
    int ntop;
    int64_t top_key[MAX_TOPS][3], top_bub[MAX_TOPS];
    int64_t bubble_id;                         /* mBubbleId */
    int nd;                                    /* the bubbles dictionary's keys in insertion order */
    int64_t dict_bub[MAX_DICTS];
    int ne;                                    /* its entries, in insertion order: (bubble, ball id, action) */
    int64_t e_bub[MAX_ENTRIES], e_id[MAX_ENTRIES];
    int e_act[MAX_ENTRIES];
    int ni, nk;                                /* bubbleInteractives and bubbleKeepAlives as (bubble, id) pairs */
    int64_t i_bub[MAX_PAIRS], i_id[MAX_PAIRS], k_bub[MAX_PAIRS], k_id[MAX_PAIRS];
    int nka;                                   /* bubbleKeepAliveBalls: ids (a removed ball's stays) */
    int64_t ka_ids[MAX_BALLS];
    OUMapBig prox;                             /* mProximityBalls */
    int32_t rnd[34];                           /* glibc's TYPE_3 table, then the front and rear indices */
    int nev;                                   /* the sensors' events: (kind, ball id, a, b) */
    int64_t ev[MAX_MEVENTS][4];
    int ngrave;                                /* removed balls' last positions: a tracking pointer outlives them */
    int64_t grave_id[MAX_GRAVE];
    double grave_pos[MAX_GRAVE][3];
    int overflow;
} OMaster;

/* ------------------------------------------------------------------ glibc rand() (random_r.c's TYPE_3 generator) */

static void m_srand(OMaster* m, uint32_t seed) {
    int32_t* r = m->rnd;
    if (seed == 0) seed = 1;
    r[0] = (int32_t)seed;
    for (int i = 1; i < 31; ++i) {             /* 16807 x mod 2^31 - 1 by Schrage's method */
        int32_t hi = r[i - 1] / 127773, lo = r[i - 1] % 127773;
        int32_t word = 16807 * lo - 2836 * hi;
        if (word < 0) word += 2147483647;
        r[i] = word;
    }
    r[31] = 3, r[32] = 0;                      /* front after the separation of 3, rear at the start */
    for (int i = 0; i < 310; ++i) {
        int32_t f = r[31], b = r[32];
        r[f] = (int32_t)((uint32_t)r[f] + (uint32_t)r[b]);
        r[31] = (f + 1) % 31, r[32] = (b + 1) % 31;
    }
}

static int32_t m_rand(OMaster* m) {
    int32_t* r = m->rnd;
    int32_t f = r[31], b = r[32];
    r[f] = (int32_t)((uint32_t)r[f] + (uint32_t)r[b]);
    int32_t out = (int32_t)((uint32_t)r[f] >> 1);
    r[31] = (f + 1) % 31, r[32] = (b + 1) % 31;
    return out;
}

/* --------------------------------------------------------- std::unordered_map<int64, ...>'s order (space.py:193-320) */

static const uint64_t PRIMES[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79,
    83, 89, 97, 103, 109, 113, 127, 137, 139, 149, 157, 167, 179, 193, 199, 211, 227, 241, 257, 277, 293, 313, 337,
    359, 383, 409, 439, 467, 503, 541, 577, 619, 661, 709, 761, 823, 887, 953, 1031};
static const uint64_t FAST_BKT[14] = {2, 2, 2, 3, 5, 5, 7, 7, 11, 11, 11, 11, 13, 13};

static uint64_t um_next_bkt(uint64_t* next, uint64_t want) {
    if (want < 14) {
        if (want == 0) return 1;
        *next = FAST_BKT[want];
        return FAST_BKT[want];
    }
    for (size_t j = 6; j < sizeof(PRIMES) / sizeof(PRIMES[0]); ++j)
        if (PRIMES[j] >= want) {
            *next = PRIMES[j];
            return PRIMES[j];
        }
    return 0;                                  /* past this oracle's table: a map above 500 keys */
}

/* the linking rule: before the first key of the order in the same bucket, else at the front */
static void um_link(int* n, int64_t* key, uint64_t bkt, int64_t k) {
    uint64_t b = (uint64_t)k % bkt;
    int at = 0;
    for (int i = 0; i < *n; ++i)
        if ((uint64_t)key[i] % bkt == b) { at = i; goto put; }
    at = 0;
put:
    memmove(&key[at + 1], &key[at], sizeof(int64_t) * (size_t)(*n - at));
    key[at] = k;
    ++*n;
}

static int um_find(int n, const int64_t* key, int64_t k) {
    for (int i = 0; i < n; ++i)
        if (key[i] == k) return i;
    return -1;
}

/* insert(k): nothing if present; else the policy asked for one more key (a growth re-links every key in order) */
static int um_insert(int* n, int64_t* key, uint64_t* bkt, uint64_t* next, int cap, int64_t k) {
    if (um_find(*n, key, k) >= 0) return 0;
    if (*n >= cap) return -1;
    uint64_t total = (uint64_t)*n + 1;
    if (total > *next) {
        uint64_t floor_keys = *next == 0 ? 11 : 0;
        double m = (double)(total > floor_keys ? total : floor_keys);
        if (m >= (double)*bkt) {
            uint64_t a = (uint64_t)floor(m) + 1, b2 = *bkt * 2;
            uint64_t nb = um_next_bkt(next, a > b2 ? a : b2);
            if (nb == 0) return -1;
            *bkt = nb;
            int64_t old[MAX_BALLS];
            int on = *n;
            memcpy(old, key, sizeof(int64_t) * (size_t)on);
            *n = 0;
            for (int i = 0; i < on; ++i) um_link(n, key, nb, old[i]);
        } else {
            *next = *bkt;
        }
    }
    um_link(n, key, *bkt, k);
    return 1;
}

static void um_erase(int* n, int64_t* key, int64_t k) {
    int i = um_find(*n, key, k);
    if (i < 0) return;
    memmove(&key[i], &key[i + 1], sizeof(int64_t) * (size_t)(*n - i - 1));
    --*n;
}

#define S_INSERT(b, k) um_insert(&(b)->s_balls.n, (b)->s_balls.key, &(b)->s_balls.bkt, &(b)->s_balls.next, 64, (k))
#define S_ERASE(b, k) um_erase(&(b)->s_balls.n, (b)->s_balls.key, (k))
#define P_INSERT(m, k) um_insert(&(m)->prox.n, (m)->prox.key, &(m)->prox.bkt, &(m)->prox.next, MAX_BALLS, (k))
#define P_ERASE(m, k) um_erase(&(m)->prox.n, (m)->prox.key, (k))

/* --------------------------------------------------------------------------------------- the top boxes */

static void top_of(const OBox* x, int64_t t[3]) {
    for (int q = 0; q < 3; ++q) t[q] = x->k[q] >> (2 * x->level);
}

static int top_index(const OMaster* m, const int64_t t[3]) {
    for (int i = 0; i < m->ntop; ++i)
        if (m->top_key[i][0] == t[0] && m->top_key[i][1] == t[1] && m->top_key[i][2] == t[2]) return i;
    return -1;
}

static int top_reached(const OPark* p, const int64_t t[3], const OBall* skip) {
    for (int i = 0; i < p->n; ++i) {
        const OBall* c = &p->balls[i];
        if (c == skip || c->level < 0) continue;
        OBox bx[8];
        int nb = ball_boxes(c, bx);
        for (int j = 0; j < nb; ++j) {
            int64_t u[3];
            top_of(&bx[j], u);
            if (u[0] == t[0] && u[1] == t[1] && u[2] == t[2]) return 1;
        }
    }
    return 0;
}

/* the table after a filing change: the tops no filing reaches released (their bubbles lost), but `pin` (a box just
   created there keeps it); the tops newly reached added with no bubble */
static void tops_sync(OPark* p, const int64_t* pin) {
    OMaster* m = p->m;
    for (int i = 0; i < m->ntop;) {
        const int64_t* t = m->top_key[i];
        int pinned = pin && t[0] == pin[0] && t[1] == pin[1] && t[2] == pin[2];
        if (!pinned && !top_reached(p, t, NULL)) {
            memmove(&m->top_key[i], &m->top_key[i + 1], sizeof(m->top_key[0]) * (size_t)(m->ntop - i - 1));
            memmove(&m->top_bub[i], &m->top_bub[i + 1], sizeof(m->top_bub[0]) * (size_t)(m->ntop - i - 1));
            --m->ntop;
        } else {
            ++i;
        }
    }
    for (int i = 0; i < p->n; ++i) {
        const OBall* c = &p->balls[i];
        if (c->level < 0) continue;
        OBox bx[8];
        int nb = ball_boxes(c, bx);
        for (int j = 0; j < nb; ++j) {
            int64_t u[3];
            top_of(&bx[j], u);
            if (top_index(m, u) < 0) {
                if (m->ntop >= MAX_TOPS) { m->overflow = 1; continue; }
                memcpy(m->top_key[m->ntop], u, sizeof(u));
                m->top_bub[m->ntop++] = -1;
            }
        }
    }
    if (pin && top_index(m, pin) < 0 && m->ntop < MAX_TOPS) {
        memcpy(m->top_key[m->ntop], pin, sizeof(int64_t) * 3);
        m->top_bub[m->ntop++] = -1;
    }
}

static int64_t top_bubble(const OMaster* m, const int64_t t[3]) {
    int i = top_index(m, t);
    return i < 0 ? -1 : m->top_bub[i];
}

static void top_set_bubble(OMaster* m, const int64_t t[3], int64_t bub) {
    int i = top_index(m, t);
    if (i >= 0) m->top_bub[i] = bub;
}

/* ------------------------------------------------------------------------------ the bubble dictionaries and sets */

static int dict_entry(const OMaster* m, int64_t bub, int64_t id) {
    for (int i = 0; i < m->ne; ++i)
        if (m->e_bub[i] == bub && m->e_id[i] == id) return i;
    return -1;
}

static void dict_ensure(OMaster* m, int64_t bub) {
    for (int i = 0; i < m->nd; ++i)
        if (m->dict_bub[i] == bub) return;
    if (m->nd >= MAX_DICTS) { m->overflow = 1; return; }
    m->dict_bub[m->nd++] = bub;
}

static int dict_exists(const OMaster* m, int64_t bub) {
    for (int i = 0; i < m->nd; ++i)
        if (m->dict_bub[i] == bub) return 1;
    return 0;
}

static void entry_del(OMaster* m, int i) {
    memmove(&m->e_bub[i], &m->e_bub[i + 1], sizeof(int64_t) * (size_t)(m->ne - i - 1));
    memmove(&m->e_id[i], &m->e_id[i + 1], sizeof(int64_t) * (size_t)(m->ne - i - 1));
    memmove(&m->e_act[i], &m->e_act[i + 1], sizeof(int) * (size_t)(m->ne - i - 1));
    --m->ne;
}

static void pair_add(int* n, int64_t* pb, int64_t* pi, int64_t bub, int64_t id, int* ovf) {
    for (int i = 0; i < *n; ++i)
        if (pb[i] == bub && pi[i] == id) return;
    if (*n >= MAX_PAIRS) { *ovf = 1; return; }
    pb[*n] = bub, pi[*n] = id;
    ++*n;
}

static void pair_del(int* n, int64_t* pb, int64_t* pi, int64_t bub, int64_t id) {
    for (int i = 0; i < *n; ++i)
        if (pb[i] == bub && pi[i] == id) {
            memmove(&pb[i], &pb[i + 1], sizeof(int64_t) * (size_t)(*n - i - 1));
            memmove(&pi[i], &pi[i + 1], sizeof(int64_t) * (size_t)(*n - i - 1));
            --*n;
            return;
        }
}

static int pair_count(int n, const int64_t* pb, int64_t bub) {
    int c = 0;
    for (int i = 0; i < n; ++i) c += pb[i] == bub;
    return c;
}

static void add_to_bubble(OPark* p, OBall* b) {                          /* space.py:922-954 */
    OMaster* m = p->m;
    if (b->bubble != -1) {
        dict_ensure(m, b->bubble);
        int e = dict_entry(m, b->bubble, b->id);
        if (e < 0) {
            if (m->ne >= MAX_ENTRIES) { m->overflow = 1; return; }
            m->e_bub[m->ne] = b->bubble, m->e_id[m->ne] = b->id, m->e_act[m->ne] = 1;
            m->ne++;
        } else if (m->e_act[e] == -1) {
            m->e_act[e] = 0;
        }
    }
    if (b->old_bubble != -1 && dict_exists(m, b->old_bubble)) {
        int e = dict_entry(m, b->old_bubble, b->id);
        if (e >= 0) {
            if (m->e_act[e] == 0) m->e_act[e] = -1;
            else if (m->e_act[e] == 1) entry_del(m, e);
        }
    }
    if (b->is_interactive) {
        if (b->bubble != -1) pair_add(&m->ni, m->i_bub, m->i_id, b->bubble, b->id, &m->overflow);
        if (b->old_bubble != -1) pair_del(&m->ni, m->i_bub, m->i_id, b->old_bubble, b->id);
    }
    int kept = 0;
    for (int i = 0; i < m->nka; ++i) kept |= m->ka_ids[i] == b->id;
    if (kept) {
        if (b->bubble != -1) pair_add(&m->nk, m->k_bub, m->k_id, b->bubble, b->id, &m->overflow);
        if (b->old_bubble != -1) pair_del(&m->nk, m->k_bub, m->k_id, b->old_bubble, b->id);
    }
}

/* InDeadBubble (space.py:974-984): not in WARP, and its bubble has no interactive and no keep-alive ball */
static int master_dead(const OPark* p, const OBall* b) {
    const OMaster* m = p->m;
    if (b->mode == WARP) return 0;
    if (pair_count(m->ni, m->i_bub, b->bubble)) return 0;
    if (pair_count(m->nk, m->k_bub, b->bubble)) return 0;
    return 1;
}

/* ---------------------------------------------------------------------------- Ballpark::InsertInBoxes on the master */

static int64_t nearby_bubbles(const OMaster* m, const int64_t t[3]) {    /* Box::NearbyBubbles, i-j-k order */
    for (int i = -1; i <= 1; ++i)
        for (int j = -1; j <= 1; ++j)
            for (int k = -1; k <= 1; ++k) {
                if (i == 0 && j == 0 && k == 0) continue;
                int64_t u[3] = {t[0] + i, t[1] + j, t[2] + k};
                int x = top_index(m, u);
                if (x >= 0 && m->top_bub[x] != -1) return m->top_bub[x];
            }
    return -1;
}

static void master_insert_ball(OPark* p, OBall* b) {                     /* space.py:729-817 */
    OMaster* m = p->m;
    const double infl = inflated_radius(p, b);
    const int level = level_for(infl);
    const double w = level_width(level);
    const int64_t n = level_grid(level);
    const double fac = 1.0 / w, half = 0.5 * w * (double)n, center = w * (double)n * 0.5;
    OBox box1 = {level, {0, 0, 0}};
    for (int q = 0; q < 3; ++q) box1.k[q] = clamp_key(to_int64_((b->pos[q] + half) * fac), n);
    int64_t top[3];
    top_of(&box1, top);
    if (top_index(m, top) < 0) tops_sync(p, top);                       /* GetBox creates box1 and its top */
    int64_t nb = top_bubble(m, top);
    if (nb == -1) {
        b->level = -1;                                                  /* DeleteFromBoxes */
        tops_sync(p, top);
        nb = nearby_bubbles(m, top);
        if (nb == -1) nb = m->bubble_id++;
        top_set_bubble(m, top, nb);
    }
    if (b->bubble != nb) {
        b->old_bubble = b->bubble;
        b->bubble = nb;
        add_to_bubble(p, b);
    }
    int mod[3];
    for (int q = 0; q < 3; ++q) {                                       /* Ball::InsertInBoxes's sides */
        double lo = (double)box1.k[q] * w - center, hi = (double)(box1.k[q] + 1) * w - center;
        mod[q] = (lo + infl - b->pos[q] > 0.0) ? -1 : (hi - infl - b->pos[q] < 0.0) ? 1 : 0;
    }
    if (b->level == level && b->key[0] == box1.k[0] && b->key[1] == box1.k[1] && b->key[2] == box1.k[2] &&
        b->mod[0] == mod[0] && b->mod[1] == mod[1] && b->mod[2] == mod[2])
        return;                                                         /* the same box and sides: nothing */
    b->level = -1;                                                      /* out of its old boxes first */
    tops_sync(p, top);
    b->level = level;
    for (int q = 0; q < 3; ++q) b->key[q] = box1.k[q], b->mod[q] = mod[q];
    tops_sync(p, top);
    top_set_bubble(m, top, nb);
    if (level == 0 && b->radius * 2 > level_width(0)) m->overflow = 1;  /* a huge ball: outside this oracle */
}

/* ------------------------------------------------------------------------------------------------ the bubbles */

static void traverse(OMaster* m, int x, int64_t seed) {                 /* Box::Traverse */
    if (m->top_bub[x] != -1) return;
    m->top_bub[x] = seed;
    int64_t t[3] = {m->top_key[x][0], m->top_key[x][1], m->top_key[x][2]};
    for (int i = -1; i <= 1; ++i)
        for (int j = -1; j <= 1; ++j)
            for (int k = -1; k <= 1; ++k) {
                if (i == 0 && j == 0 && k == 0) continue;
                int64_t u[3] = {t[0] + i, t[1] + j, t[2] + k};
                int y = top_index(m, u);
                if (y >= 0) traverse(m, y, seed);
            }
}

static int by_seq(const OPark* p, int* idx) {                           /* the ball dictionary's order */
    for (int i = 0; i < p->n; ++i) idx[i] = i;
    for (int i = 1; i < p->n; ++i) {
        int v = idx[i], k = i;
        while (k > 0 && p->balls[idx[k - 1]].seq > p->balls[v].seq) { idx[k] = idx[k - 1]; --k; }
        idx[k] = v;
    }
    return p->n;
}

static void initialize_bubbles(OPark* p) {                              /* space.py:993-1015 */
    OMaster* m = p->m;
    for (int i = 0; i < m->ntop; ++i) m->top_bub[i] = -1;
    m->bubble_id = 0;
    m->nd = m->ne = m->ni = m->nk = m->nka = 0;
    int idx[MAX_BALLS];
    int n = by_seq(p, idx);
    for (int q = 0; q < n; ++q) {
        OBall* b = &p->balls[idx[q]];
        if (b->level < 0) {
            b->old_bubble = b->bubble, b->bubble = -1;
            continue;
        }
        OBox my = {b->level, {b->key[0], b->key[1], b->key[2]}};
        int64_t t[3];
        top_of(&my, t);
        int x = top_index(m, t);
        if (m->top_bub[x] != -1) {
            b->old_bubble = b->bubble, b->bubble = m->top_bub[x];
            add_to_bubble(p, b);
            continue;
        }
        traverse(m, x, m->bubble_id);
        b->old_bubble = b->bubble, b->bubble = m->bubble_id;
        add_to_bubble(p, b);
        m->bubble_id++;
    }
}

static void update_ball_bubble(OPark* p, int64_t id) {                  /* space.py:1017-1033 */
    OBall* ball = find(p, id);
    if (!ball) return;
    const double max_width = level_width(0) * level_width(0);
    int64_t hi_id = -1;
    int idx[MAX_BALLS];
    int n = by_seq(p, idx);
    for (int q = 0; q < n; ++q) {
        OBall* b = &p->balls[idx[q]];
        if (b->bubble > hi_id) hi_id = b->bubble;
        if (b == ball || b->bubble == -1) continue;
        double d[3] = {ball->pos[0] - b->pos[0], ball->pos[1] - b->pos[1], ball->pos[2] - b->pos[2]};
        if (lensq(d) < max_width) {
            ball->bubble = b->bubble;
            break;
        }
    }
    if (ball->bubble == -1) ball->bubble = hi_id + 1;
}

static void copy_bubbles(OMaster* m) {                                  /* space.py:1035-1051 */
    for (int d = 0; d < m->nd;) {
        const int64_t bub = m->dict_bub[d];
        for (int i = 0; i < m->ne;) {
            if (m->e_bub[i] != bub) { ++i; continue; }
            if (m->e_act[i] == 1) m->e_act[i] = 0;
            if (m->e_act[i] == -1) { entry_del(m, i); continue; }
            ++i;
        }
        if (!pair_count(m->ne, m->e_bub, bub)) {
            memmove(&m->dict_bub[d], &m->dict_bub[d + 1], sizeof(int64_t) * (size_t)(m->nd - d - 1));
            --m->nd;
        } else {
            ++d;
        }
    }
}

/* ----------------------------------------------------------------------------------------------- the sensors */

static void m_event(OMaster* m, int kind, int64_t id, int64_t a, int64_t b) {
    if (m->nev >= MAX_MEVENTS) { m->overflow = 1; return; }
    m->ev[m->nev][0] = kind, m->ev[m->nev][1] = id, m->ev[m->nev][2] = a, m->ev[m->nev][3] = b;
    m->nev++;
}

static void update_proximity_status(OPark* p, OBall* b) {               /* Ball::UpdateProximityStatus */
    if (b->trk_range != 0.0 || b->ntf_range != 0.0 || b->sensor) {
        b->has_prox = 1;
        if (P_INSERT(p->m, b->id) < 0) p->m->overflow = 1;
    } else {
        b->has_prox = 0;
        P_ERASE(p->m, b->id);
    }
}

/* the master's RemoveBall at once (space.py:1286-1317, after the stop and the release): DoFinalCleanup, out of its
   boxes and its bubble, out of the proximity map; its position kept for the tracking pointers that outlive it */
static void master_removed(OPark* p, OBall* b) {
    OMaster* m = p->m;
    m_event(m, EV_FINAL_CLEANUP, b->id, 0, 0);
    b->level = -1;
    tops_sync(p, NULL);
    b->old_bubble = b->bubble, b->bubble = -1;
    add_to_bubble(p, b);
    if (b->has_prox) P_ERASE(m, b->id);
    if (m->ngrave < MAX_GRAVE) {
        m->grave_id[m->ngrave] = b->id;
        memcpy(m->grave_pos[m->ngrave++], b->pos, sizeof(b->pos));
    }
}

static const double* position_of(const OPark* p, int64_t id) {
    const OBall* b = find((OPark*)p, id);
    if (b) return b->pos;
    for (int i = p->m->ngrave - 1; i >= 0; --i)
        if (p->m->grave_id[i] == id) return p->m->grave_pos[i];
    return NULL;
}

/* GetNearbyBalls for a sensor (exclude_fixed unless cloaked, exclude_inert when it senses interactives only; no
   missile, cloak or harmonic filter): the walk of oracle_modes.c's box set, the master's bubble rule */
static int sensor_candidates(OPark* p, int bi, int exclude_fixed, int exclude_inert, int* out) {
    const OBall* b = &p->balls[bi];
    if (exclude_inert && pair_count(p->m->ni, p->m->i_bub, b->bubble) == 0) return 0;
    OBoxes* U = &g_boxes;
    boxes_build(p, U);
    OBox bx[8];
    int nb = ball_boxes(b, bx);
    int sbox[MAX_BALLS * 8 * NUM_LEVELS];
    int ns = 0;
    for (int j = 0; j < nb; ++j) {
        int bj = box_index(U, bx[j]);
        if (U->handled[bj]) continue;
        OBox x = bx[j];
        while (x.level > 0) {
            x = parent_of(x);
            int xi = box_index(U, x);
            if (U->handled[xi]) break;
            if (box_ball_count(p, x) > 0) {
                sbox[ns++] = xi;
                U->handled[xi] = 1;
            }
        }
        explode_children(p, U, bx[j], sbox, &ns);
        if (box_ball_count(p, bx[j]) > 1) {
            sbox[ns++] = bj;
            U->handled[bj] = 1;
        }
    }
    int handled[MAX_BALLS] = {0};
    int cnt = 0;
    for (int s = 0; s < ns; ++s) {
        OBox x = U->box[sbox[s]];
        U->handled[sbox[s]] = 0;
        for (int i = 0; i < p->n; ++i) {                                /* the box's balls by id */
            const OBall* o = &p->balls[i];
            if (!ball_in_box(o, x) || handled[i] || i == bi || o->is_moribund) continue;
            if (exclude_inert) {
                if (!o->is_interactive) continue;
            } else {
                if (o->bubble != b->bubble) continue;                   /* no global balls here */
                if (exclude_fixed && !o->is_free) continue;
            }
            handled[i] = 1;
            out[cnt++] = i;
        }
    }
    return cnt;
}

static void set_radius_refile(OPark* p, OBall* b, double r) {           /* SetBallRadius(Ball*, float) */
    if (r < 0.0) return;
    b->radius = r;
    master_insert_ball(p, b);
}

static void check_sensor(OPark* p, OBall* b) {                          /* space.py:1476-1522 */
    OMaster* m = p->m;
    b->s_elapsed += p->dt;
    if (b->s_elapsed < b->s_period) return;
    b->s_elapsed = 0.0;
    const double radius = b->radius;
    const double vel[3] = {b->vel[0], b->vel[1], b->vel[2]};
    b->vel[0] = b->vel[1] = b->vel[2] = 0.0;
    set_radius_refile(p, b, c_float(b->radius + b->s_range));
    const int64_t me = b->id;
    int uni[MAX_BALLS];
    int nu = sensor_candidates(p, (int)(b - p->balls), !b->cloaked, b->s_only, uni);
    int64_t uni_id[MAX_BALLS];
    for (int i = 0; i < nu; ++i) uni_id[i] = p->balls[uni[i]].id;
    b = find(p, me);
    for (int k = 0; k < 3; ++k) b->vel[k] = vel[k];
    set_radius_refile(p, b, radius);
    if (b->s_balls.n > 0) {
        int64_t keys[64];
        int nk = b->s_balls.n;
        memcpy(keys, b->s_balls.key, sizeof(int64_t) * (size_t)nk);
        for (int i = 0; i < nk; ++i) {
            int near = keys[i] == me;
            for (int j = 0; j < nu && !near; ++j) near = uni_id[j] == keys[i];
            if (near) continue;
            S_ERASE(b, keys[i]);
            m_event(m, EV_PROXIMITY, me, keys[i], 0);
            if (!b->sensor) break;
        }
    }
    for (int j = 0; j < nu; ++j) {
        if (b->is_moribund) break;
        OBall* o = find(p, uni_id[j]);
        if (!o || o->id == me || o->is_moribund) continue;
        const int in = um_find(b->s_balls.n, b->s_balls.key, o->id) >= 0;
        double d[3] = {b->pos[0] - o->pos[0], b->pos[1] - o->pos[1], b->pos[2] - o->pos[2]};
        const double d2 = lensq(d);
        const double ck = c_float(c_float(b->s_range + o->radius) + b->radius);
        if (d2 > ck * ck) {
            if (in) {
                S_ERASE(b, o->id);
                m_event(m, EV_PROXIMITY, me, o->id, 0);
            }
        } else if (!in) {
            if (S_INSERT(b, o->id) < 0) m->overflow = 1;
            m_event(m, EV_PROXIMITY, me, o->id, 1);
        }
        if (!b->sensor) break;
    }
}

static void check_notification(OPark* p, OBall* b) {                    /* space.py:1434-1458 */
    int send = 0, inside = 0;
    if (b->ntf_range != 0.0) {
        const double nr = b->ntf_range;
        if (b->mode == GOTO) {
            const double lim = nr + b->radius;
            double d[3] = {b->pos[0] - b->gp[0], b->pos[1] - b->gp[1], b->pos[2] - b->gp[2]};
            inside = lensq(d) < lim * lim;
        } else if (b->mode == FOLLOW || b->mode == MISSILE || b->mode == ORBIT) {
            const OBall* o = find(p, b->follow_ptr);
            double d[3] = {b->pos[0] - o->pos[0], b->pos[1] - o->pos[1], b->pos[2] - o->pos[2]};
            const double len2 = lensq(d);
            const double hi = o->radius + b->follow_range + nr + b->radius;
            double lo = hi - 2.0 * nr;
            if (lo < 0.0) lo = 0.0;
            inside = len2 < hi * hi && len2 > lo * lo;
        }
        if (inside) {
            if (!b->ntf_in) send = 1, b->ntf_in = 1;
        } else if (b->ntf_in) {
            send = 1, b->ntf_in = 0;
        }
    }
    if (send) m_event(p->m, EV_RANGE, b->id, b->ntf_in, 0);
}

static void check_tracking(OPark* p, OBall* b) {                        /* space.py:1460-1474 */
    int send = 0;
    if (b->trk_range >= 0.00001 && b->trk_id != 0) {
        const double* tp = position_of(p, b->trk_id);
        double d[3] = {b->pos[0] - tp[0], b->pos[1] - tp[1], b->pos[2] - tp[2]};
        const double d2 = lensq(d), lim = b->trk_range + b->radius;
        if (d2 < lim * lim) {
            if (!b->trk_in) send = 1, b->trk_in = 1;
        } else if (b->trk_in) {
            send = 1, b->trk_in = 0;
        }
    }
    if (send) m_event(p->m, EV_TRACKING, b->id, b->trk_in, 0);
}

static void check_for_proximity(OPark* p, OBall* b) {                   /* Ball::CheckForProximity */
    if (b->is_moribund) return;
    check_notification(p, b);
    check_tracking(p, b);
    if (!b->sensor) return;
    check_sensor(p, b);
}

static void master_proximities(OPark* p) {                              /* space.py:1406-1422 */
    OMaster* m = p->m;
    int64_t keys[MAX_BALLS];
    const int nk = m->prox.n;
    memcpy(keys, m->prox.key, sizeof(int64_t) * (size_t)nk);
    for (int i = 0; i < nk; ++i) {
        OBall* b = find(p, keys[i]);
        if (!b || b->is_moribund) continue;
        if (master_dead(p, b)) {
            if (b->s_balls.n > 0) {
                b->s_elapsed = b->s_period;
                check_for_proximity(p, b);
                b = find(p, keys[i]);
            }
            b->s_balls.n = 0;                                           /* clear(): the buckets kept */
            continue;
        }
        if (!b->has_prox) continue;
        check_for_proximity(p, b);
    }
}

/* ------------------------------------------------------------------------------------------------- the API */

void or_park_set_master(OPark* p, uint32_t seed) {
    if (!p->m) p->m = (OMaster*)calloc(1, sizeof(OMaster));
    p->m->prox.bkt = 1;
    m_srand(p->m, seed);
}

void or_park_free_master(OPark* p) { free(p->m); p->m = NULL; }

int or_master_overflow(const OPark* p) { return p->m ? p->m->overflow : 0; }

/* The master's operations: 1 keep alive (bubbleKeepAliveBalls.add), 2 AddProximitySensor(range rr, period x, shuffle
   w & 1, interactives only w & 2), 3 RemoveProximitySensor, 4 SetNotificationRange(rr), 5 SetTargetTracking(target
   other, rr), 6 SetBallInteractive(w), 7 InitializeBubbles (id 0) or UpdateBallBubble(id), 8 CopyBubbles. */
void or_master_op(OPark* p, int op, int64_t id, int64_t other, double rr, double x, int64_t w) {
    OMaster* m = p->m;
    OBall* b = find(p, id);
    switch (op) {
        case 1: {
            int have = 0;
            for (int i = 0; i < m->nka; ++i) have |= m->ka_ids[i] == id;
            if (!have && m->nka < MAX_BALLS) m->ka_ids[m->nka++] = id;
            break;
        }
        case 2: {
            if (!b) break;
            const double range = c_float(rr);
            if (c_float(b->radius + range) < 0.0) break;
            b->s_balls.n = 0;
            b->s_range = range;
            b->s_period = x;
            b->sensor = 1;
            b->s_only = (w & 2) != 0;
            if (w & 1) b->s_elapsed = 0.0 + (double)m_rand(m) / 2147483647.0 * (b->s_period - 0.0);
            update_proximity_status(p, b);
            break;
        }
        case 3:
            if (b) b->sensor = 0, update_proximity_status(p, b);
            break;
        case 4:
            if (!b || rr < 0.0) break;
            b->ntf_range = rr;
            update_proximity_status(p, b);
            b->ntf_in = 0;
            break;
        case 5:
            if (!b) break;
            if (rr <= 0.0 || other == 0) {
                b->trk_id = 0, b->trk_range = 0.0;
                break;
            }
            if (!find(p, other)) break;
            b->trk_id = other, b->trk_range = rr;
            update_proximity_status(p, b);
            b->trk_in = 0;
            break;
        case 6:
            if (!b) break;
            if (w && !b->is_interactive) {
                b->is_interactive = 1;
                if (b->bubble != -1) pair_add(&m->ni, m->i_bub, m->i_id, b->bubble, b->id, &m->overflow);
            } else if (!w && b->is_interactive) {
                if (b->bubble != -1) pair_del(&m->ni, m->i_bub, m->i_id, b->bubble, b->id);
                b->is_interactive = 0;
            }
            break;
        case 7:
            if (id == 0) initialize_bubbles(p);
            else update_ball_bubble(p, id);
            break;
        case 8: copy_bubbles(m); break;
        default: break;
    }
}

/* Readers. Per ball in ascending id: bubble, old bubble (2 a ball). */
int or_master_bubbles(const OPark* p, int64_t* out) {
    for (int i = 0; i < p->n; ++i) out[2 * i] = p->balls[i].bubble, out[2 * i + 1] = p->balls[i].old_bubble;
    return p->n;
}

/* The tops: key (3), bubble; returns their count. */
int or_master_tops(const OPark* p, int64_t* out) {
    for (int i = 0; i < p->m->ntop; ++i)
        for (int q = 0; q < 4; ++q) out[4 * i + q] = q < 3 ? p->m->top_key[i][q] : p->m->top_bub[i];
    return p->m->ntop;
}

/* The dictionaries in their order: per dictionary its bubble, its entry count, then (id, action) pairs; returns the
   words written. */
int or_master_dicts(const OPark* p, int64_t* out) {
    const OMaster* m = p->m;
    int k = 0;
    out[k++] = m->nd;
    for (int d = 0; d < m->nd; ++d) {
        out[k++] = m->dict_bub[d];
        const int at = k++;
        int c = 0;
        for (int i = 0; i < m->ne; ++i)
            if (m->e_bub[i] == m->dict_bub[d]) out[k++] = m->e_id[i], out[k++] = m->e_act[i], ++c;
        out[at] = c;
    }
    return k;
}

/* which 0: the interactive pairs, 1 the keep-alive pairs (bubble, id); returns the count. */
int or_master_pairs(const OPark* p, int which, int64_t* out) {
    const OMaster* m = p->m;
    const int n = which ? m->nk : m->ni;
    for (int i = 0; i < n; ++i) {
        out[2 * i] = which ? m->k_bub[i] : m->i_bub[i];
        out[2 * i + 1] = which ? m->k_id[i] : m->i_id[i];
    }
    return n;
}

/* The proximity map (k < 0) or ball k's sensor list: keys in order, then its bucket count and next resize; returns
   the key count. */
int or_master_umap(const OPark* p, int k, int64_t* out) {
    if (k < 0) {
        const OUMapBig* u = &p->m->prox;
        memcpy(out, u->key, sizeof(int64_t) * (size_t)u->n);
        out[u->n] = (int64_t)u->bkt, out[u->n + 1] = (int64_t)u->next;
        return u->n;
    }
    const OBall* b = &p->balls[k];
    memcpy(out, b->s_balls.key, sizeof(int64_t) * (size_t)b->s_balls.n);
    out[b->s_balls.n] = (int64_t)b->s_balls.bkt, out[b->s_balls.n + 1] = (int64_t)b->s_balls.next;
    return b->s_balls.n;
}

int64_t or_master_bubble_id(const OPark* p) { return p->m->bubble_id; }

int or_master_keep_balls(const OPark* p, int64_t* out) {
    memcpy(out, p->m->ka_ids, sizeof(int64_t) * (size_t)p->m->nka);
    return p->m->nka;
}

/* The events since the last call (kind, id, a, b), cleared; returns their count. */
int or_master_events(OPark* p, int64_t* out, int cap) {
    int n = p->m->nev < cap ? p->m->nev : cap;
    memcpy(out, p->m->ev, sizeof(int64_t) * 4 * (size_t)n);
    p->m->nev = 0;
    return n;
}

int64_t or_master_visibility(OPark* p, int64_t src_id, int64_t dst_id) {    /* space.py:1524-1548 */
    const OBall* src = find(p, src_id);
    const OBall* dst = find(p, dst_id);
    if (!src || !dst) return -2;
    if (src->bubble != dst->bubble) return -2;
    double ray[3] = {dst->pos[0] - src->pos[0], dst->pos[1] - src->pos[1], dst->pos[2] - src->pos[2]};
    const double distance = sqrt(lensq(ray));
    double u[3];
    normalized(ray, u);
    int idx[MAX_BALLS];
    int n = by_seq(p, idx);
    for (int q = 0; q < n; ++q) {
        const OBall* occ = &p->balls[idx[q]];
        if (occ->bubble != src->bubble || occ == src || occ == dst || !occ->is_massive || occ->cloaked) continue;
        double tc[3] = {occ->pos[0] - src->pos[0], occ->pos[1] - src->pos[1], occ->pos[2] - src->pos[2]};
        const double v = u[0] * tc[0] + u[1] * tc[1] + u[2] * tc[2];
        const double r = occ->radius;
        const double dsq = r * r - ((tc[0] * tc[0] + tc[1] * tc[1] + tc[2] * tc[2]) - v * v);
        if (dsq < 0.0) continue;
        const double d = sqrt(dsq);
        if (v - d > 0.0 && v - d < distance) return occ->id;
    }
    return 0;
}

static int64_t existing_top_bubble(const OPark* p, const double pos[3], int* found) {
    const int level = level_for(10000000.0);
    const double w = level_width(level);
    const int64_t n = level_grid(level);
    const double fac = 1.0 / w, half = 0.5 * w * (double)n;
    int64_t t[3];
    for (int q = 0; q < 3; ++q) t[q] = clamp_key(to_int64_((pos[q] + half) * fac), n);
    int x = top_index(p->m, t);
    *found = x >= 0;
    return x >= 0 ? p->m->top_bub[x] : -1;
}

/* GetBallIdsInRange (and its distances): from ball `id` when by_ball, else from (x, y, z); the bubble dictionary's
   balls in its order within range plus their radius; returns the count, or -1 for Destiny's error (no such ball). */
int or_master_range(OPark* p, int by_ball, int64_t id, double x, double y, double z, double range, int cloaked,
                    int64_t* out, double* d2_out) {
    const OBall* ref = NULL;
    double c[3] = {x, y, z};
    int64_t bub;
    if (by_ball) {
        ref = find(p, id);
        if (!ref) return -1;
        memcpy(c, ref->pos, sizeof(c));
        bub = ref->bubble;
    } else {
        int found;
        bub = existing_top_bubble(p, c, &found);
        if (!found) bub = -1;
    }
    if (bub == -1 || !dict_exists(p->m, bub)) return 0;
    int n = 0;
    const OMaster* m = p->m;
    for (int i = 0; i < m->ne; ++i) {
        if (m->e_bub[i] != bub) continue;
        const OBall* o = find(p, m->e_id[i]);
        if (!o || o == ref) continue;
        if (o->is_moribund) continue;
        if (o->cloaked && !cloaked) continue;
        double d[3] = {o->pos[0] - c[0], o->pos[1] - c[1], o->pos[2] - c[2]};
        const double d2 = lensq(d), lim = range + o->radius;
        if (d2 <= lim * lim) {
            out[n] = o->id;
            d2_out[n++] = d2;
        }
    }
    return n;
}

int64_t or_master_bubble_at(const OPark* p, double x, double y, double z, int* found) {
    double c[3] = {x, y, z};
    return existing_top_bubble(p, c, found);
}

int or_master_remote(OPark* p, int64_t* out) {                          /* space.py:1062-1072 */
    int n = 0;
    for (int i = 0; i < p->n; ++i) {
        const OBall* src = &p->balls[i];
        if (!src->is_free) continue;
        if (src->mode == FOLLOW || src->mode == ORBIT || src->mode == MISSILE || src->mode == FORMATION) {
            const OBall* dst = find(p, src->follow_id);
            if (dst && src->bubble == dst->bubble) continue;
            out[n++] = src->id;
        }
    }
    return n;
}
