/*
 * clothmi.hpp — C++ simulation-object wrapper over the C ABI of clothmi.h.
 *
 * Shaped like the reference's physics-body interface (IPhysicsbody,
 * Physics/IPhysicsbody.h:7-20) and its GPU cloth (Physics::CudaDeformableBody,
 * Physics/CudaDeformableBody.h:22-24): the constructor creates the body
 * (CudaDeformableBody.cpp:30-285), FixedUpdate runs one substep
 * (CudaDeformableBody.cpp:550-627), Update reads back positions and vertex normals
 * (CudaDeformableBody.cpp:629-638), SetSpherePos moves the sphere (:640-643).
 * Header-only; it depends on clothmi.h and the C++ standard library only, so a
 * caller maps its own vertex type (the reference's VertexFormat) onto the flat
 * arrays. Errors throw std::runtime_error carrying cm_last_error().
 */
#ifndef CLOTHMI_HPP
#define CLOTHMI_HPP

#include <stdexcept>
#include <string>
#include <vector>

#include "clothmi.h"

namespace clothmi {

/* The interface shape of IPhysicsbody (Physics/IPhysicsbody.h:7-20). */
class PhysicsBody {
 public:
  virtual ~PhysicsBody() {}
  virtual void FixedUpdate() = 0;
  virtual void Update() = 0;
  virtual const std::vector<unsigned int>& GetTriangles() const = 0;
};

class ClothBodyHip : public PhysicsBody {
 public:
  /* positions: 3N floats; indices: 3T triangle vertex ids; fixed: N flags (pinned
   * vertices, inverse mass 0); inv_mass: inverse mass of every free vertex. */
  ClothBodyHip(const std::vector<float>& positions, const std::vector<unsigned int>& indices,
               const cm_params& params, const std::vector<bool>& fixed, float inv_mass)
      : m_indices(indices) {
    const int n = (int)(positions.size() / 3);
    if (positions.size() != 3 * (size_t)n || fixed.size() != (size_t)n || indices.size() % 3)
      throw std::invalid_argument("clothmi: inconsistent array sizes");
    std::vector<uint8_t> pinned(n);
    std::vector<float> w(n);
    for (int i = 0; i < n; ++i) {
      pinned[i] = fixed[i] ? 1 : 0;
      w[i] = fixed[i] ? 0.f : inv_mass;
    }
    check(cm_create(&params, positions.data(), n, indices.data(), (int)(indices.size() / 3),
                    pinned.data(), w.data(), nullptr, 0, nullptr, nullptr, &m_body));
    m_positions = positions;
    m_normals.assign(positions.size(), 0.f);
  }
  ~ClothBodyHip() override { cm_destroy(m_body); }
  ClothBodyHip(const ClothBodyHip&) = delete;
  ClothBodyHip& operator=(const ClothBodyHip&) = delete;

  /* One substep on the body's stream; does not wait for it. */
  void FixedUpdate() override { check(cm_step(m_body, 1)); }

  /* Waits for the substeps in flight and reads positions and vertex normals. */
  void Update() override {
    check(cm_read_positions(m_body, m_positions.data()));
    check(cm_read_normals(m_body, m_normals.data()));
  }

  /* Update into the caller's Rendering::VertexFormat array (N records of 48 bytes):
   * positions and normals written, color and uv kept (cm_read_vertices). */
  void UpdateVertices(void* vertex_format_records) {
    check(cm_read_vertices(m_body, static_cast<float*>(vertex_format_records)));
  }

  void SetSpherePos(const float center[3], float radius) {
    check(cm_set_sphere(m_body, center, radius));
  }

  const std::vector<unsigned int>& GetTriangles() const override { return m_indices; }
  const std::vector<float>& Positions() const { return m_positions; }  // after Update
  const std::vector<float>& Normals() const { return m_normals; }      // after Update
  cm_body* Handle() const { return m_body; }

 private:
  static void check(int rc) {
    if (rc != CM_OK) throw std::runtime_error(std::string("clothmi: ") + cm_last_error());
  }
  cm_body* m_body = nullptr;
  std::vector<unsigned int> m_indices;
  std::vector<float> m_positions, m_normals;
};

}  // namespace clothmi

#endif /* CLOTHMI_HPP */
