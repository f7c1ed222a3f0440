// Prints the GlobalData block the reference writes each frame, computed with the
// reference's own vendored glm 0.9.9.7, for camera states given on stdin
// ("px py pz rx ry rz" per line).  The calls reproduce
// src/engine/rendering/materials.cpp:532-545; output is one JSON array per line
// of 76 floats (camPos, aspect, camDir, view, proj, viewproj, invViewProj).
#define GLM_ENABLE_EXPERIMENTAL
#include <glm/glm.hpp>
#include <glm/gtx/transform.hpp>

#include <cstdio>

int main() {
  float p[3], r[3];
  while (std::scanf("%f %f %f %f %f %f", &p[0], &p[1], &p[2], &r[0], &r[1], &r[2]) == 6) {
    const glm::vec3 camPos(p[0], p[1], p[2]), camRot(r[0], r[1], r[2]);
    glm::mat4 view = glm::rotate(glm::mat4{1.0f}, glm::radians(0.0f), glm::vec3(1, 0, 0)) *
                     glm::rotate(glm::mat4{1.0f}, glm::radians(camRot.y), glm::vec3(0, 1, 0)) *
                     glm::translate(glm::mat4(1.0f), camPos);
    glm::mat4 proj = glm::perspective(glm::radians(70.0f), 1700.0f / 900.0f, 0.01f, 200.0f);
    proj[1][1] *= -1;
    glm::mat4 viewproj = proj * view;
    glm::mat4 inv = glm::inverse(viewproj);
    float out[76];
    const float head[12] = {camPos.x, camPos.y, camPos.z, 1.0f, 1700.0f / 900.0f, 850, 450, 0, 0, 0, 1, 1};
    for (int i = 0; i < 12; i++) out[i] = head[i];
    const glm::mat4 *mats[4] = {&view, &proj, &viewproj, &inv};
    for (int m = 0; m < 4; m++)
      for (int c = 0; c < 4; c++)
        for (int k = 0; k < 4; k++) out[12 + 16 * m + 4 * c + k] = (*mats[m])[c][k];
    std::printf("[");
    for (int i = 0; i < 76; i++) std::printf("%s%.9g", i ? "," : "", out[i]);
    std::printf("]\n");
  }
  return 0;
}
