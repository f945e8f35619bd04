// Geometry metrics (reference include/System/Metrics.h, src/System/Metrics.cpp:
// 12-82): distance, angle = acos(clamp(cos)) (NaN if an arm is degenerate),
// dihedral = atan2(|b2| b1.n2, n1.n2) (NaN if degenerate).
#ifndef MD_AMD_METRICS_H
#define MD_AMD_METRICS_H

namespace Metrics {
double distance(double x1, double y1, double z1, double x2, double y2, double z2);
double angle(double x1, double y1, double z1, double x2, double y2, double z2,
             double x3, double y3, double z3);
double dihedral(double x1, double y1, double z1, double x2, double y2, double z2,
                double x3, double y3, double z3, double x4, double y4, double z4);
}  // namespace Metrics

#endif
