/* Oracle (test infrastructure): CPU restatement of the reference's run-time
 * command scheduler, used only by tests to check the product's scheduler.
 *
 *   libSteve/SimCommand.c  scmdaddcommand :273-336, scmdstr2cmd :340-380,
 *                          scmddocommandtiming :384-420,
 *                          scmdupdatecommands :423-477, scmdpop :480-491,
 *                          scmdexecute :494-545
 *   libSteve/queue.c       q_insert :204-256 (order: key, then list position)
 *   Smoldyn/smolsim.cpp    simdocommands :2572-2582, simulatetimestep's time
 *                          handling :2680-2690, smolsimulate :2747-2760
 *   Smoldyn/libsmoldyn.cpp smolRunTimeStep :217-222 (break time t+dt/2),
 *                          smolRunSimUntil :274-282 (break time = stop time after)
 *
 * SimCommand.c defines SMOLDYN in its own text and then includes
 * smoldynconfigure.h, which CMake generates, so the reference scheduler does
 * not compile on its own here.  This restatement is pinned by the reference's
 * tests (tests/test_data_in_array.py: an 'E' command over 1000 steps gives
 * 1001 rows starting at time 0; tests/test_data_outputfile.py: 'i 0 100 10'
 * at dt 0.1 gives 11 rows whose times average 50.036) and by closed forms of
 * each timing code in tests/test_commands.py.
 *
 * Script lines: "times tmin tmax dt", "cmd T on off step mult text",
 * "cmdstr line", "dt v", "tmax v", "run", "runstep", "rununtil t".  A command
 * whose text starts with "stop" returns CMDstop.  Events: (time, list
 * position, iteration) per execution, (time, -1000-code, iteration) per run.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OC_MAXCMD 64
#define OC_STR 256

enum { OK = 0, WARN = 1, PAUSE = 2, STOP = 3, ABORT = 4 };

typedef struct {
  char timing;
  double on, off, dt, xt;
  long long oni, offi, dti;
  char str[OC_STR];
} ocmd;

typedef struct {
  double kd;
  long long kl;
  int ki;     /* list position */
  ocmd c;
} qitem;

typedef struct {
  int n;
  qitem it[OC_MAXCMD];
} oqueue;

typedef struct {
  int condition;
  long long iter;
  int ncmd;
  ocmd tmpl[OC_MAXCMD];
  int twin[OC_MAXCMD];    /* a copy is queued */
  int invoke[OC_MAXCMD];  /* a copy was made since the last reset */
  oqueue fq, iq;
  /* clock */
  double tmin, tmax, dt, time, tbreak;
  /* events */
  int nev, maxev;
  double* et;
  int* ec;
  long long* ei;
} ostate;

/* q_insert: items with a larger key, or an equal key and larger position,
 * move back one place */
static void q_put(oqueue* q, double kd, long long kl, int isfloat, int ki, const ocmd* c) {
  int i = q->n;
  while (i > 0) {
    const qitem* p = &q->it[i - 1];
    int later = isfloat ? (kd < p->kd || (kd == p->kd && ki < p->ki)) : (kl < p->kl || (kl == p->kl && ki < p->ki));
    if (!later) break;
    q->it[i] = q->it[i - 1];
    i--;
  }
  q->it[i].kd = kd;
  q->it[i].kl = kl;
  q->it[i].ki = ki;
  q->it[i].c = *c;
  q->n++;
}

static qitem q_take(oqueue* q) {
  qitem x = q->it[0];
  memmove(q->it, q->it + 1, sizeof(qitem) * (size_t)(q->n - 1));
  q->n--;
  return x;
}

static void setcondition(ostate* s, int cond, int upgrade) {
  if (upgrade == 0 && s->condition > cond) s->condition = cond;
  else if (upgrade == 1 && s->condition < cond) s->condition = cond;
  else if (upgrade == 2) s->condition = cond;
}

static int addcommand(ostate* s, char timing, double on, double off, double step, double mult, const char* str) {
  ocmd c;
  memset(&c, 0, sizeof(c));
  if (!str || !*str) return 3;
  c.timing = timing;
  if (strchr("baBAEe", timing) && timing) {
  } else if (timing == '@') {
    c.on = c.off = on;
  } else if (timing == 'i') {
    if (step <= 0) return 5;
    c.on = on;
    c.off = off;
    c.dt = step;
  } else if (timing == 'x') {
    if (step <= 0) return 5;
    c.on = on;
    c.off = off;
    c.dt = step;
    c.xt = mult;
  } else if (timing == '&') {
    if ((double)(long long)on != on) return 4;
    c.oni = c.offi = (long long)on;
    c.dti = 1;
  } else if (timing == 'I' || timing == 'j') {
    if ((double)(long long)on != on || (double)(long long)off != off || (double)(long long)step != step) return 4;
    if (step <= 0) return 5;
    c.oni = (long long)on;
    c.offi = (long long)off;
    c.dti = (long long)step;
  } else if (timing == 'N' || timing == 'n') {
    if ((double)(long long)step != step) return 4;
    if (step <= 0) return 5; /* the reference loops without end for a zero period */
    c.dti = (long long)step;
  } else
    return 6;
  strncpy(c.str, str, OC_STR - 1);
  if (c.str[0] && c.str[strlen(c.str) - 1] == '\n') c.str[strlen(c.str) - 1] = 0;
  if (s->ncmd >= OC_MAXCMD) return 1;
  s->tmpl[s->ncmd] = c;
  s->twin[s->ncmd] = 0;
  s->invoke[s->ncmd] = 0;
  s->ncmd++;
  setcondition(s, 2, 0);
  return 0;
}

/* strnword-style: pointer to the word after the first n words */
static const char* skipwords(const char* p, int n) {
  for (int k = 0; k < n; k++) {
    while (*p == ' ' || *p == '\t') p++;
    if (!*p) return NULL;
    while (*p && *p != ' ' && *p != '\t') p++;
  }
  while (*p == ' ' || *p == '\t') p++;
  return *p ? p : NULL;
}

static int str2cmd(ostate* s, const char* line) {
  double v[4] = {0, 0, 0, 0};
  char timing;
  const char* rest;
  while (*line == ' ' || *line == '\t') line++;
  if (!*line) return 3;
  timing = line[0];
  rest = skipwords(line, 1);
  if (!rest) return 3;
  if (strchr("baBAEe", timing)) {
  } else if (timing == '@' || timing == '&') {
    if (sscanf(rest, "%lf", &v[0]) != 1) return 3;
    rest = skipwords(rest, 1);
  } else if (timing == 'N' || timing == 'n') {
    if (sscanf(rest, "%lf", &v[2]) != 1) return 3;
    rest = skipwords(rest, 1);
  } else if (timing == 'i' || timing == 'I' || timing == 'j') {
    if (sscanf(rest, "%lf %lf %lf", &v[0], &v[1], &v[2]) != 3) return 3;
    rest = skipwords(rest, 3);
  } else if (timing == 'x') {
    if (sscanf(rest, "%lf %lf %lf %lf", &v[0], &v[1], &v[2], &v[3]) != 4) return 3;
    rest = skipwords(rest, 4);
  } else
    return 6;
  if (!rest) return 3;
  return addcommand(s, timing, v[0], v[1], v[2], v[3], rest);
}

static void docommandtiming(ocmd* c, double tmin, double tmax, double dt, long long iter) {
  long long nst = (long long)((tmax - tmin) / dt + 0.5);
  char t = c->timing;
  if (t == 'b') {
    c->on = c->off = tmin - dt;
    c->dt = dt;
  } else if (t == 'a') {
    c->on = c->off = tmax + dt;
    c->dt = dt;
  } else if (t == '@')
    c->dt = dt;
  else if (t == 'i' || t == 'x') {
    if (c->on < tmin) c->on = tmin;
    if (c->off > tmax) c->off = tmax;
  } else if (t == 'B') {
    c->oni = c->offi = iter - 1;
    c->dti = 1;
  } else if (t == 'A') {
    c->oni = c->offi = iter + nst + 1;
    c->dti = 1;
  } else if (t == 'I' || t == 'j') {
    if (c->oni < 0) c->oni = iter + 1;
  } else if (t == 'E' || t == 'e') {
    c->oni = iter;
    c->offi = iter + nst;
    c->dti = 1;
  } else if (t == 'N' || t == 'n') {
    c->oni = iter;
    c->offi = iter + nst;
  }
}

static void updatecommands(ostate* s) {
  if (s->condition == 3) return;
  if (s->dt <= 0 || s->tmax < s->tmin) return;
  if (s->condition == 0) {
    s->fq.n = s->iq.n = 0;
    for (int i = 0; i < s->ncmd; i++) s->twin[i] = s->invoke[i] = 0;
  }
  for (int i = 0; i < s->ncmd; i++) {
    if (!s->twin[i] && !s->invoke[i]) {
      ocmd c = s->tmpl[i];
      docommandtiming(&c, s->tmin, s->tmax, s->dt, s->iter);
      if (strchr("ba@ix", c.timing)) q_put(&s->fq, c.on, 0, 1, i, &c);
      else q_put(&s->iq, 0, c.oni, 0, i, &c);
      s->twin[i] = s->invoke[i] = 1;
    } else if (s->condition == 1 && s->twin[i]) {
      for (int k = 0; k < s->fq.n; k++)
        if (s->fq.it[k].ki == i) docommandtiming(&s->fq.it[k].c, s->tmin, s->tmax, s->dt, s->iter);
      for (int k = 0; k < s->iq.n; k++)
        if (s->iq.it[k].ki == i) docommandtiming(&s->iq.it[k].c, s->tmin, s->tmax, s->dt, s->iter);
    }
  }
  setcondition(s, 3, 1);
}

static void event(ostate* s, double t, int c, long long it) {
  if (s->nev < s->maxev) {
    if (s->et) s->et[s->nev] = t;
    if (s->ec) s->ec[s->nev] = c;
    if (s->ei) s->ei[s->nev] = it;
  }
  s->nev++;
}

static int runone(ostate* s, ocmd* c, long long iter) {
  int pos = -1;
  for (int i = 0; i < s->ncmd; i++)
    if (!strcmp(s->tmpl[i].str, c->str)) {
      pos = i;
      break;
    }
  event(s, s->time, pos, iter);
  return strncmp(c->str, "stop", 4) == 0 ? STOP : OK;
}

static int execute(ostate* s, int donow) {
  int code2 = OK, code1;
  long long iter = s->iter++;
  while (s->iq.n > 0 && (s->iq.it[0].kl <= iter || donow)) {
    qitem x = q_take(&s->iq);
    code1 = runone(s, &x.c, iter);
    if (x.c.oni + x.c.dti <= x.c.offi && !donow && (code1 == OK || code1 == PAUSE)) {
      x.c.oni += x.c.dti;
      q_put(&s->iq, 0, x.c.oni, 0, x.ki, &x.c);
    } else
      s->twin[x.ki] = 0;
    if (code1 == ABORT) return code1;
    if (code1 > code2) code2 = code1;
  }
  while (s->fq.n > 0 && (s->fq.it[0].kd <= s->time || donow)) {
    qitem x = q_take(&s->fq);
    double dt;
    code1 = runone(s, &x.c, iter);
    dt = x.c.dt >= s->dt ? x.c.dt : s->dt;
    if (x.c.on + dt <= x.c.off && !donow && (code1 == OK || code1 == PAUSE)) {
      x.c.on += dt;
      if (x.c.xt > 1) x.c.dt *= x.c.xt;
      q_put(&s->fq, x.c.on, 0, 1, x.ki, &x.c);
    } else
      s->twin[x.ki] = 0;
    if (code1 == ABORT) return code1;
    if (code1 > code2) code2 = code1;
  }
  return code2;
}

static void cmdpop(ostate* s, double t) {
  while (s->fq.n > 0 && s->fq.it[0].kd <= t) {
    qitem x = q_take(&s->fq);
    s->twin[x.ki] = 0;
  }
}

static int simulate(ostate* s) {
  int er = 0, code;
  updatecommands(s);
  code = execute(s, 0);
  if (code == STOP || code == ABORT) er = 7;
  while (!er) {
    s->time += s->dt;
    code = execute(s, 0);
    if (code == STOP || code == ABORT) er = 7;
    else if (s->time >= s->tmax) er = 1;
    else if (s->time >= s->tbreak) er = 10;
  }
  if (er != 10) {
    cmdpop(s, s->tmax);
    execute(s, 1);
    setcondition(s, 0, 0);
  }
  event(s, s->time, -1000 - er, s->iter);
  return er;
}

int ocmd_timeline(const char* script, int maxev, double* et, int* ec, long long* ei) {
  ostate* s = (ostate*)calloc(1, sizeof(ostate));
  char line[1024];
  const char* p = script;
  int nev;
  if (!s || !script) {
    free(s);
    return -1;
  }
  s->tmax = 100;
  s->dt = 0.01;
  s->tbreak = 1e300;
  s->maxev = maxev;
  s->et = et;
  s->ec = ec;
  s->ei = ei;
  while (*p) {
    const char* e = strchr(p, '\n');
    size_t len = e ? (size_t)(e - p) : strlen(p);
    char word[32];
    int nr = 0;
    if (len >= sizeof(line)) len = sizeof(line) - 1;
    memcpy(line, p, len);
    line[len] = 0;
    p = e ? e + 1 : p + len;
    if (sscanf(line, "%31s%n", word, &nr) != 1) continue;
    if (!strcmp(word, "times")) {
      if (sscanf(line + nr, "%lf %lf %lf", &s->tmin, &s->tmax, &s->dt) != 3) goto bad;
      s->time = s->tmin;
      setcondition(s, 1, 0);
    } else if (!strcmp(word, "cmd")) {
      char t;
      double on, off, st, mu;
      int n2 = 0;
      if (sscanf(line + nr, " %c %lf %lf %lf %lf %n", &t, &on, &off, &st, &mu, &n2) != 5) goto bad;
      addcommand(s, t, on, off, st, mu, line + nr + n2);
    } else if (!strcmp(word, "cmdstr")) {
      str2cmd(s, line + nr);
    } else if (!strcmp(word, "dt")) {
      if (sscanf(line + nr, "%lf", &s->dt) != 1) goto bad;
      setcondition(s, 1, 0);
    } else if (!strcmp(word, "tmax")) {
      if (sscanf(line + nr, "%lf", &s->tmax) != 1) goto bad;
    } else if (!strcmp(word, "run")) {
      simulate(s);
    } else if (!strcmp(word, "runstep")) {
      s->tbreak = s->time + s->dt / 2;
      simulate(s);
    } else if (!strcmp(word, "rununtil")) {
      double t, stop = s->tmax;
      if (sscanf(line + nr, "%lf", &t) != 1) goto bad;
      s->tbreak = t;
      simulate(s);
      s->tbreak = stop;
    } else
      goto bad;
  }
  nev = s->nev;
  free(s);
  return nev;
bad:
  free(s);
  return -1;
}
