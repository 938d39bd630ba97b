/*
 * Tags (reference src/core/dependencies/tags.c): a task with use_tag = 1
 * carries tag `tag_id`; starpu_tag_declare_deps(id, ...) makes the task
 * carrying `id` wait for the tasks carrying the listed tags;
 * starpu_tag_wait blocks until a tag's task has completed.
 * STARPU_TAG_ONLY (task insert) only names the task: no dependency.
 */
#define _GNU_SOURCE
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>

#include "sp_internal.h"

struct sp_tag {
	starpu_tag_t id;
	int done;
	int ndeps;               /* unfinished tags this one depends on */
	struct sp_job *job;      /* task carrying the tag, if submitted */
	struct sp_tag **succ;
	int nsucc, cap;
	int used;
};

static unsigned long hash_tag(starpu_tag_t id)
{
	uint64_t z = id + 0x9E3779B97F4A7C15ull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return (unsigned long)(z ^ (z >> 31));
}

static struct sp_tag *tag_lookup(starpu_tag_t id, int create)
{
	if (sp.tags_cap == 0) {
		if (!create)
			return NULL;
		sp.tags_cap = 1024;
		sp.tags = calloc((size_t)sp.tags_cap, sizeof(*sp.tags));
	}
	unsigned long mask = (unsigned long)sp.tags_cap - 1;
	unsigned long i = hash_tag(id) & mask;
	while (sp.tags[i]) {
		if (sp.tags[i]->id == id)
			return sp.tags[i];
		i = (i + 1) & mask;
	}
	if (!create)
		return NULL;
	if (2 * (sp.ntags + 1) > sp.tags_cap) {
		/* grow and rehash */
		struct sp_tag **old = sp.tags;
		int oldcap = sp.tags_cap;
		sp.tags_cap *= 2;
		sp.tags = calloc((size_t)sp.tags_cap, sizeof(*sp.tags));
		mask = (unsigned long)sp.tags_cap - 1;
		for (int k = 0; k < oldcap; k++)
			if (old[k]) {
				unsigned long p = hash_tag(old[k]->id) & mask;
				while (sp.tags[p])
					p = (p + 1) & mask;
				sp.tags[p] = old[k];
			}
		free(old);
		i = hash_tag(id) & mask;
		while (sp.tags[i])
			i = (i + 1) & mask;
	}
	struct sp_tag *t = calloc(1, sizeof(*t));
	t->id = id;
	sp.tags[i] = t;
	sp.ntags++;
	return t;
}

static void tag_release_job(struct sp_tag *t)
{
	struct sp_job *j = t->job;
	if (j && j->tag_pending) {
		j->tag_pending = 0;
		if (--j->ndeps == 0 && j->submitted)
			sp_job_ready(j);
	}
}

static void tag_mark_done(struct sp_tag *t)
{
	t->done = 1;
	for (int i = 0; i < t->nsucc; i++) {
		struct sp_tag *s = t->succ[i];
		if (--s->ndeps == 0)
			tag_release_job(s);
	}
	t->nsucc = 0;
	pthread_cond_broadcast(&sp.done_cond);
}

void starpu_tag_declare_deps_array(starpu_tag_t id, unsigned ndeps, starpu_tag_t *array)
{
	SP_LOCK();
	struct sp_tag *t = tag_lookup(id, 1);
	for (unsigned i = 0; i < ndeps; i++) {
		struct sp_tag *d = tag_lookup(array[i], 1);
		if (d->done)
			continue;
		if (d->nsucc == d->cap) {
			d->cap = d->cap ? 2 * d->cap : 4;
			d->succ = realloc(d->succ, (size_t)d->cap * sizeof(*d->succ));
		}
		d->succ[d->nsucc++] = t;
		t->ndeps++;
		if (t->job && t->job->submitted && !t->job->tag_pending && !t->job->running && !t->job->terminated) {
			/* deps declared after submission are not supported (as in the reference) */
		}
	}
	SP_UNLOCK();
}

void starpu_tag_declare_deps(starpu_tag_t id, unsigned ndeps, ...)
{
	starpu_tag_t arr[256];
	STARPU_ASSERT(ndeps <= 256);
	va_list ap;
	va_start(ap, ndeps);
	for (unsigned i = 0; i < ndeps; i++)
		arr[i] = va_arg(ap, starpu_tag_t);
	va_end(ap);
	starpu_tag_declare_deps_array(id, ndeps, arr);
}

void sp_tag_job_submitted(struct sp_job *j)
{
	struct starpu_task *t = j->task;
	if (!t || !t->use_tag)
		return;
	struct sp_tag *tg = tag_lookup(t->tag_id, 1);
	tg->job = j;
	tg->done = 0;
	if (tg->ndeps > 0) {
		j->ndeps++;
		j->tag_pending = 1;
	}
}

void sp_tag_job_done(struct sp_job *j)
{
	struct starpu_task *t = j->task;
	if (!t || !t->use_tag)
		return;
	struct sp_tag *tg = tag_lookup(t->tag_id, 1);
	tg->job = NULL;
	tag_mark_done(tg);
}

int starpu_tag_wait(starpu_tag_t id)
{
	SP_LOCK();
	struct sp_tag *t = tag_lookup(id, 1);
	while (!t->done)
		pthread_cond_wait(&sp.done_cond, &sp.lock);
	SP_UNLOCK();
	return 0;
}

int starpu_tag_wait_array(unsigned ntags, starpu_tag_t *id)
{
	for (unsigned i = 0; i < ntags; i++)
		starpu_tag_wait(id[i]);
	return 0;
}

void starpu_tag_notify_from_apps(starpu_tag_t id)
{
	SP_LOCK();
	struct sp_tag *t = tag_lookup(id, 1);
	if (!t->done)
		tag_mark_done(t);
	SP_UNLOCK();
}

void starpu_tag_restart(starpu_tag_t id)
{
	SP_LOCK();
	struct sp_tag *t = tag_lookup(id, 0);
	if (t)
		t->done = 0;
	SP_UNLOCK();
}

void starpu_tag_remove(starpu_tag_t id)
{
	SP_LOCK();
	if (sp.tags_cap) {
		unsigned long mask = (unsigned long)sp.tags_cap - 1;
		unsigned long i = hash_tag(id) & mask;
		while (sp.tags[i] && sp.tags[i]->id != id)
			i = (i + 1) & mask;
		if (sp.tags[i]) {
			/* mark as removed but keep the slot chain intact */
			sp.tags[i]->done = 1;
			sp.tags[i]->nsucc = 0;
		}
	}
	SP_UNLOCK();
}

void sp_tags_free(void)
{
	for (int i = 0; i < sp.tags_cap; i++)
		if (sp.tags[i]) {
			free(sp.tags[i]->succ);
			free(sp.tags[i]);
		}
	free(sp.tags);
	sp.tags = NULL;
	sp.tags_cap = sp.ntags = 0;
}
