/*
 * Host build of the engine's binary32 culling test (csrc/cull.h) for
 * tests/test_culling.py: is one box pair culled at one query?
 */
#include <string.h>

#include "../../puzzletunneldiscovery_amd/csrc/cull.h"

using namespace se3cq;

static GNode node_of(const float* v)
{
	GNode g;
	g.c_ex = {v[0], v[1], v[2], v[3]};
	g.a0_ey = {v[4], v[5], v[6], v[7]};
	g.a1_ez = {v[8], v[9], v[10], v[11]};
	g.a2_child = {v[12], v[13], v[14], v[15]};
	return g;
}

/* env: GNode fields of an environment box (c, ex, axis0, ey, axis1, ez,
 * axis2, child); rob: the same for a robot box in the robot's frame; q: the
 * query state; pert: the perturbation (null: identity).  margins = 0 drops
 * the culling's inflation and translation margins (for checking that the
 * test notices).  Returns 1 when the pair is culled. */
extern "C" int cc_culled(const float* env, const float* rob, const double* q, const double* pert, float rob_reach,
                         float scene_mag, int margins)
{
	DevScene s;
	memset(&s, 0, sizeof(s));
	PertData P;
	if (pert) {
		P = make_pert(pert);
		s.pert = &P;
		s.pert_identity = 0;
	} else {
		s.pert_identity = 1;
	}
	s.rob_reach = rob_reach;
	s.scene_mag = scene_mag;
	QCull c;
	make_qcull(s, q, c);
	if (!margins) {
		c.infl = 0.0f;
		c.eta = 0.0f;
	}
	RobBox B;
	load_rob_box(node_of(rob), c, B);
	return obb_disjoint(node_of(env), B, c.eta) ? 1 : 0;
}
